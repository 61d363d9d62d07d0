"""Host likelihood layer (milkywayathome_client_amd/likelihood.py) against
the reference's own known answers and closed forms:

* nbody/tests/EMD_Range_test.c: its two histogram files (tests/golden/
  EMD_range_{1,2}.hist, copied as fixtures) give nbMatchEMD = 0 with the
  first file's lambda ranges and 5.449191188301354 over the whole histogram;
* nbody/tests/emd_test.c's distributions on (lambda, beta) grids: point
  masses in opposite corners or two random bins (EMD = their Euclidean
  distance), one shared bin and a random histogram against itself (0);
* nbCostComponent worked by hand, the file round trip and the worst-case /
  mismatch conventions of nbSystemLikelihood."""
import math
import os

import numpy as np
import pytest

from milkywayathome_client_amd import histogram as H
from milkywayathome_client_amd import likelihood as L

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_emd_range_known_answers():
    h1 = L.read_histogram(os.path.join(GOLDEN, "EMD_range_1.hist"))
    h2 = L.read_histogram(os.path.join(GOLDEN, "EMD_range_2.hist"))
    assert h1.emd_range == [-33.0, -27.0, -24.0, -18.0]
    assert L.match_emd(h1, h2) == 0.0                    # scoreWithRange
    assert abs(L.match_emd(h2, h1) - 5.449191188301354) < 1e-12   # scoreWithoutRange


def _grid(d1, d2):
    lam, beta = np.meshgrid(np.arange(d1, dtype=float), np.arange(d2, dtype=float),
                            indexing="ij")
    return lam.ravel(), beta.ravel()


@pytest.mark.parametrize("d1,d2", [(1, 7), (7, 1), (7, 7), (11, 11), (11, 4)])
def test_emd_test_distributions(d1, d2):
    rng = np.random.default_rng(d1 * 100 + d2)
    lam, beta = _grid(d1, d2)
    n = lam.size
    # oppositeSides
    w1, w2 = np.zeros(n), np.zeros(n)
    w1[0], w2[n - 1] = 1.0, 1.0
    expect = math.hypot(lam[0] - lam[-1], beta[0] - beta[-1])
    assert L.emd_calc(w1, lam, beta, w2, lam, beta) == pytest.approx(expect, abs=1e-9)
    # allInDifferentBins
    for _ in range(3):
        i, j = rng.integers(0, n, size=2)
        w1, w2 = np.zeros(n), np.zeros(n)
        w1[i], w2[j] = 1.0, 1.0
        expect = math.hypot(lam[i] - lam[j], beta[i] - beta[j])
        assert L.emd_calc(w1, lam, beta, w2, lam, beta) == pytest.approx(expect, abs=1e-9)
    # allInSameBin and randomSelf
    i = rng.integers(0, n)
    w1 = np.zeros(n)
    w1[i] = 1.0
    assert L.emd_calc(w1, lam, beta, w1, lam, beta) == pytest.approx(0.0, abs=1e-12)
    w = rng.uniform(size=n)
    w /= w.sum()
    assert L.emd_calc(w, lam, beta, w, lam, beta) == pytest.approx(0.0, abs=1e-9)


def test_one_dimensional_emd_equals_linear_program():
    """The closed 1-D form against the general transport solution (the same
    points moved off a common beta by a negligible amount force the LP)."""
    rng = np.random.default_rng(3)
    lam = np.linspace(-150, 150, 50)
    beta = np.zeros(50)
    w1 = rng.uniform(size=50)
    w2 = rng.uniform(size=50)
    w1 /= w1.sum()
    w2 /= w2.sum()
    exact = L.emd_calc(w1, lam, beta, w2, lam, beta)
    lp = L.emd_calc(w1, lam, beta, w2, lam, beta + 1e-12)
    assert exact == pytest.approx(lp, rel=1e-9)
    # shifting a distribution by one bin costs one bin width
    w = np.zeros(50)
    w[10] = 1.0
    ws = np.roll(w, 1)
    assert L.emd_calc(w, lam, beta, ws, lam, beta) == pytest.approx(lam[1] - lam[0])


def test_two_dimensional_emd_on_a_50_by_20_grid():
    """A 50 x 20 lambda-beta grid (10^6 flows, sparse constraints): moving
    all weight one beta row costs one row spacing; a histogram against
    itself costs 0.  Larger problems are refused, not attempted."""
    lam_c = np.linspace(-150, 150, 50)
    beta_c = np.linspace(-15, 15, 20)
    lam, beta = (a.ravel() for a in np.meshgrid(lam_c, beta_c, indexing="ij"))
    rng = np.random.default_rng(5)
    w = np.zeros((50, 20))
    w[:, 7] = rng.uniform(0.5, 1.0, size=50)
    w /= w.sum()
    assert L.emd_calc(w.ravel(), lam, beta, np.roll(w, 1, axis=1).ravel(), lam, beta) == \
        pytest.approx(beta_c[1] - beta_c[0], rel=1e-9)
    big = np.ones(2001)
    with pytest.raises(ValueError, match="transport variables"):
        L.emd_calc(big, np.arange(2001.0), np.arange(2001.0), big, np.arange(2001.0),
                   np.arange(2001.0))


def _hist(var, err, n, tot, mass, use=None):
    nb = len(var)
    return L.Histogram(lambda_bins=nb, beta_bins=1, total_num=n, total_simulated=tot,
                       mass_per_particle=mass,
                       use_bin=np.ones(nb, dtype=np.int64) if use is None else np.array(use),
                       lam=np.arange(nb, dtype=float), beta=np.zeros(nb),
                       variable=np.array(var, dtype=float), err=np.array(err, dtype=float),
                       emd_range=[])


def test_cost_component_by_hand():
    data = _hist([0.5, 0.5], [0.05, 0.05], 100, 100, 1.0)
    sim = _hist([0.25, 0.75], [0.0, 0.0], 80, 100, 1.0)
    # p = 80/100; var_data = 2 (0.05 * 100)^2 = 50;
    # cost = (100 - 80)^2 / (2 (50 + 80 * 0.8 * 0.2)) = 400 / 125.6
    assert L.cost_component(data, sim) == pytest.approx(400.0 / 125.6, rel=1e-14)
    # a bin left out of the comparison removes its simulated bodies
    data_cut = _hist([0.5, 0.5], [0.05, 0.05], 100, 100, 1.0, use=[1, 0])
    n_sim = 80 - 60                                  # round(0.75 * 80) dropped
    p = n_sim / 100.0
    expect = (100.0 - n_sim) ** 2 / (2.0 * (25.0 + n_sim * p * (1.0 - p)))
    assert L.cost_component(data_cut, sim) == pytest.approx(expect, rel=1e-14)
    # conventions: empty -> inf, no mass -> NaN, mismatched bins -> NaN
    assert L.cost_component(data, _hist([0.25, 0.75], [0, 0], 0, 100, 1.0)) == math.inf
    assert math.isnan(L.cost_component(data, _hist([0.25, 0.75], [0, 0], 80, 100, 0.0)))
    assert math.isnan(L.cost_component(data, _hist([1.0], [0.0], 80, 100, 1.0)))


def test_identical_histograms_have_zero_likelihood():
    h = _hist([0.1, 0.2, 0.3, 0.4], [0.01] * 4, 1000, 1000, 2e-4)
    h.emd_range = [0.0, 3.0]
    out = L.system_likelihood(h, h, 1000)
    assert out[0] == 0.0 and out[1] == 0.0 and out[2] == 0.0
    assert np.all(np.isnan(out[3:]))


def test_worst_case_and_mismatch():
    data = _hist([0.5, 0.5], [0.05, 0.05], 100, 100, 1.0)
    few = _hist([0.5, 0.5], [0.0, 0.0], 2, 100, 1.0)
    out = L.system_likelihood(data, few, nbody=100000)  # 2 < 1e-4 * 1e5
    assert out[0] == out[1] == L.DEFAULT_WORST_CASE and np.all(out[2:] == 0.0)
    assert np.all(np.isnan(L.system_likelihood(data, _hist([1.0], [0.0], 50, 100, 1.0), 100)))
    for flag in ("use_beta_disp", "use_vel_disp", "use_vlos", "use_beta_comp", "use_dist",
                 "use_prop_mot"):
        with pytest.raises(ValueError):                  # no per-bin columns
            L.system_likelihood(data, data, 100, **{flag: True})
    # an EMD above 50 (the maximum distance) is NaN
    far = _hist([1.0, 0.0], [0.0, 0.0], 100, 100, 1.0)
    near = _hist([0.0, 1.0], [0.0, 0.0], 100, 100, 1.0)
    far.lam = near.lam = np.array([0.0, 60.0])
    far.emd_range = near.emd_range = [0.0, 60.0]
    assert math.isnan(L.match_emd(far, near))


def test_from_counts_file_round_trip(tmp_path):
    hp = H.HistogramParams()
    rng = np.random.default_rng(5)
    pos = rng.normal(size=(20000, 3)) * 10.0 + np.array([-20.0, 5.0, 30.0])
    raw = np.zeros(hp.lambdaBins, dtype=np.int64)
    lam, beta = H.lambda_beta(pos, hp)
    li = np.floor((lam - hp.lambdaStart) / ((hp.lambdaEnd - hp.lambdaStart) / hp.lambdaBins))
    bi = np.floor((beta - hp.betaStart) / (hp.betaEnd - hp.betaStart))
    ok = (li >= 0) & (li < hp.lambdaBins) & (bi == 0)
    np.add.at(raw, li[ok].astype(np.int64), 1)
    h = L.from_counts(raw, hp, total_simulated=20000, mass_per_particle=1e-3)
    ref, _ = H.create_histogram(pos, hp)
    np.testing.assert_allclose(h.variable, ref, rtol=1e-15, atol=0)
    assert h.lam[0] == hp.lambdaStart + 0.5 * (hp.lambdaEnd - hp.lambdaStart) / hp.lambdaBins
    k = int(np.argmax(raw))
    c = float(raw[k])
    assert h.err[k] == pytest.approx(math.sqrt(c * (c / 20000) * (1 - c / 20000)) / raw.sum())
    path = tmp_path / "out.hist"
    L.write_histogram(str(path), h, comment="round trip")
    back = L.read_histogram(str(path))
    np.testing.assert_allclose(back.variable, h.variable, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(back.lam, h.lam, rtol=1e-14)
    assert (back.total_num, back.total_simulated) == (h.total_num, h.total_simulated)
    assert back.mass_per_particle == pytest.approx(h.mass_per_particle, rel=1e-12)
    assert L.match_emd(back, h) == 0.0


def test_nineteen_column_round_trip(tmp_path):
    """With histograms 1-7 the file carries 'hasPM = 1', the '!' line and 19
    columns (nbPrintHistogram); reading it back gives every column, and a
    reader of the 15-column format ignores the proper-motion columns, as the
    reference's sscanf does."""
    rng = np.random.default_rng(6)
    h = _hist(rng.uniform(size=6), rng.uniform(size=6) * 0.01, 600, 1000, 1e-3)
    for k in range(1, 8):
        h.comps[k] = (rng.normal(size=6), rng.uniform(size=6))
    path = tmp_path / "pm.hist"
    L.write_histogram(str(path), h)
    text = path.read_text()
    assert "hasPM = 1\n!\n" in text
    back = L.read_histogram(str(path))
    for k in range(1, 8):
        np.testing.assert_allclose(back.comps[k][0], h.comps[k][0], rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(back.comps[k][1], h.comps[k][1], rtol=1e-14, atol=1e-15)
    path.write_text(text.replace("hasPM = 1\n", ""))
    back = L.read_histogram(str(path))
    assert sorted(back.comps) == [1, 2, 3, 4, 5]
    np.testing.assert_allclose(back.comps[5][0], h.comps[5][0], rtol=1e-14, atol=1e-15)
