"""The best-likelihood run of nbRunSystemPlain (milkywayathome_client_amd/
plain.py, nbody/src/nbody_plain.c:75-245,366-461) on the CPU, over a stand-in
state whose histogram at each step is known: the steps evaluated (from the
BestLikeStart fraction on), the improvements counted, the time of the best,
the clamp of the likelihood (NaN and overflow -> worst case, 0 -> best
case) and the components left at 0 when not in use."""
import math
from types import SimpleNamespace

import numpy as np
import pytest

from milkywayathome_client_amd import histogram as H
from milkywayathome_client_amd import likelihood as L
from milkywayathome_client_amd import plain as P


class _State:
    """Steps a counter; at step s the histogram is the target moved by
    |s - target_step| bins (so the EMD grows with the distance in steps)."""

    def __init__(self, n_step, target_step, hp, time_evolve=2.0):
        self.ctx = SimpleNamespace(nStep=n_step, timeEvolve=time_evolve)
        self.n = 1000
        self.s = 0
        self.target = target_step
        self.hp = hp
        self.gravity_calls = 0
        self.base = np.zeros(hp.lambdaBins, dtype=np.uint32)
        self.base[20:25] = [50, 100, 200, 100, 50]

    def gravity(self):
        self.gravity_calls += 1

    def step(self):
        self.s += 1

    def tree_info(self):
        return SimpleNamespace(step=self.s)

    def counts_at(self, s):
        return np.roll(self.base, abs(s - self.target))

    def histogram(self, hp):
        c = self.counts_at(self.s)
        return c[None, :], np.array([c.sum()], dtype=np.uint32)

    def histogram_dispersion(self, hp, sim=0, **kw):
        nb = hp.lambdaBins * hp.betaBins
        return {k: (np.full(nb, 1.0 + 0.1 * abs(self.s - self.target)), np.full(nb, 0.1))
                for k in range(1, 8)}

    def bodies(self):
        return (np.full((2, 3), float(self.s)),) * 3


def test_best_step_time_and_count():
    hp = H.HistogramParams()
    st = _State(20, 14, hp)
    data = L.from_counts(st.counts_at(14), hp, 1000, 1e-3)
    best = P.run_best_likelihood(st, data, hp, 1000, 1e-3, best_like_start=0.5,
                                 keep_bodies=True, record=True)
    assert st.gravity_calls == 1 and st.s == 20
    assert [s for s, _ in best.history] == list(range(10, 21))
    # improvements at steps 10, 11, 12, 13 and 14; the identical histogram
    # has likelihood 0, reported as the best case
    assert best.step == 14 and best.count == 5
    assert best.likelihood == P.DEFAULT_BEST_CASE
    assert best.time == pytest.approx(14 / 20 * 2.0, rel=0, abs=0)
    assert best.emd == 0.0 and best.mass == 0.0
    assert best.beta == 0.0 and best.pm_dec == 0.0               # terms not in use
    assert best.bodies[0][0, 0] == 14.0
    np.testing.assert_array_equal(best.histogram.variable, data.variable)
    likes = dict(best.history)
    assert likes[12] > likes[13] > likes[14] and likes[15] == likes[13]   # one bin either way


def test_components_in_use_and_start_fraction():
    hp = H.HistogramParams()
    st = _State(10, 3, hp)
    data = L.from_counts(st.counts_at(3), hp, 1000, 1e-3,
                         dispersions=st.histogram_dispersion(hp))
    best = P.run_best_likelihood(st, data, hp, 1000, 1e-3, best_like_start=0.95,
                                 use_beta_disp=True, use_prop_mot=True, record=True)
    # 9 / 10 < 0.95: only the last step is evaluated
    assert [s for s, _ in best.history] == [10]
    assert best.count == 1 and best.step == 10
    assert best.beta > 0.0 and best.pm_dec > 0.0 and best.pm_ra > 0.0
    assert best.likelihood == pytest.approx(best.emd + best.mass + best.beta + best.pm_dec
                                            + best.pm_ra, rel=1e-14)


def test_clamp():
    assert P.clamp_likelihood(math.nan) == L.DEFAULT_WORST_CASE
    assert P.clamp_likelihood(2 * L.DEFAULT_WORST_CASE) == L.DEFAULT_WORST_CASE
    assert P.clamp_likelihood(-2 * L.DEFAULT_WORST_CASE) == L.DEFAULT_WORST_CASE
    assert P.clamp_likelihood(0.0) == P.DEFAULT_BEST_CASE
    assert P.clamp_likelihood(3.5) == 3.5 and P.clamp_likelihood(-3.5) == -3.5


def test_worse_than_worst_never_improves():
    hp = H.HistogramParams()
    st = _State(4, 2, hp)
    # a data histogram with other bins: every step's likelihood is NaN
    # (mismatched bin counts) -> clamped to the worst case, never better
    other = L.from_counts(np.ones(10, dtype=np.uint32), H.HistogramParams(lambdaBins=10),
                          1000, 1e-3)
    best = P.run_best_likelihood(st, other, hp, 1000, 1e-3, best_like_start=0.0, record=True)
    assert best.count == 0 and best.step == -1
    assert best.likelihood == L.DEFAULT_WORST_CASE
    assert all(v == L.DEFAULT_WORST_CASE for _, v in best.history)
