"""Host-side histogram and EMD (milkywayathome_client_amd/histogram.py):
lambda/beta against the frame's defining property, binning against numpy's
histogram2d, the EMD against scipy's 1-D Wasserstein distance and its
closed form for a shifted delta."""
import math

import numpy as np
import pytest
from scipy.stats import wasserstein_distance

from milkywayathome_client_amd import histogram as H


def _points(hp, lam_deg, beta_deg, dist=20.0):
    R = hp.rotation()
    out = []
    for l, b in zip(np.radians(lam_deg), np.radians(beta_deg)):
        # the frame flips the sign of the third axis before asin (beta = asin(-z/r))
        t = dist * np.array([math.cos(b) * math.cos(l), math.cos(b) * math.sin(l), -math.sin(b)])
        out.append(R.T @ t - np.array([hp.sunGCDist, 0.0, 0.0]))
    return np.array(out)


def test_rotation_is_orthonormal_and_frame_round_trips():
    hp = H.HistogramParams()
    R = hp.rotation()
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-15)
    lam0 = np.array([-140.0, -10.0, 0.0, 33.3, 149.0])
    beta0 = np.array([-12.0, 0.0, 5.5, 14.0, -3.0])
    lam, beta = H.lambda_beta(_points(hp, lam0, beta0), hp)
    np.testing.assert_allclose(lam, lam0, atol=1e-11)
    np.testing.assert_allclose(beta, beta0, atol=1e-11)


def test_binning_matches_histogram2d():
    hp = H.HistogramParams(betaBins=3)
    rng = np.random.default_rng(5)
    lam0 = rng.uniform(-170, 170, size=4000)
    beta0 = rng.uniform(-20, 20, size=4000)
    counts, total = H.create_histogram(_points(hp, lam0, beta0), hp)
    ref, _, _ = np.histogram2d(lam0, beta0, bins=(hp.lambdaBins, hp.betaBins),
                               range=((hp.lambdaStart, hp.lambdaEnd), (hp.betaStart, hp.betaEnd)))
    assert total == int(ref.sum())
    np.testing.assert_allclose(counts, ref.ravel() / ref.sum(), atol=1e-14)


def test_ignored_bodies_are_left_out():
    hp = H.HistogramParams()
    pts = _points(hp, [0.0, 10.0, 20.0], [0.0, 0.0, 0.0])
    counts, total = H.create_histogram(pts, hp, ignore=[False, True, False])
    assert total == 2 and counts.sum() == 1.0


def test_emd_against_scipy_and_closed_form():
    hp = H.HistogramParams()
    rng = np.random.default_rng(6)
    a = rng.uniform(size=50)
    b = rng.uniform(size=50)
    a /= a.sum()
    b /= b.sum()
    w = (hp.lambdaEnd - hp.lambdaStart) / hp.lambdaBins
    centres = hp.lambdaStart + w * (np.arange(50) + 0.5)
    ref = wasserstein_distance(centres, centres, a, b)
    assert abs(H.emd(a, b, hp) - ref) < 2e-9
    d1 = np.zeros(50)
    d1[10] = 1.0
    d2 = np.zeros(50)
    d2[17] = 1.0
    assert abs(H.emd(d1, d2, hp) - 7 * w) < 1e-9
    assert H.emd(a, a, hp) == 0.0


def test_emd_over_several_beta_bins():
    """With several beta bins the EMD is the transport distance between bin
    centres: a unit mass moved from one bin to another costs their Euclidean
    distance in (lambda, beta) degrees; a histogram against itself costs 0;
    one beta bin keeps the closed form."""
    hp = H.HistogramParams(betaBins=3)
    nb_ = hp.lambdaBins * hp.betaBins
    a = np.zeros(nb_)
    b = np.zeros(nb_)
    a[10 * 3 + 0] = 1.0           # lambda bin 10, beta bin 0
    b[13 * 3 + 2] = 1.0           # lambda bin 13, beta bin 2
    ls = (hp.lambdaEnd - hp.lambdaStart) / hp.lambdaBins
    bs = (hp.betaEnd - hp.betaStart) / hp.betaBins
    assert H.emd(a, b, hp) == pytest.approx(np.hypot(3 * ls, 2 * bs), abs=1e-8)
    rng = np.random.default_rng(2)
    w = rng.uniform(size=nb_)
    assert H.emd(w, w, hp) == pytest.approx(0.0, abs=1e-8)
    one = H.HistogramParams()
    x = np.zeros(one.lambdaBins)
    y = np.zeros(one.lambdaBins)
    x[5], y[9] = 1.0, 1.0
    assert H.emd(x, y, one) == pytest.approx(4 * (one.lambdaEnd - one.lambdaStart)
                                             / one.lambdaBins, abs=1e-9)
