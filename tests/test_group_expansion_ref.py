"""The far-field expansion restated in oracle/group_ref.py (the checker of
nbhip_set_group_expansion) against the reference interaction it
approximates: a(D - u) ~ g - J u + K[u,u]/2 about a group's centre.  Its
error must fall as the cube of |u|/|D| for a monopole (g, J and K right)
and at least as the square with a quadrupole (g and J of the quadrupole
right; its K is left out), and vanish at u = 0."""
import numpy as np
import pytest

from oracle import group_ref as gr


def _errors(m, Q, ratios, seed=1):
    rng = np.random.default_rng(seed)
    D = np.array([3.0, -2.0, 4.0])
    u0 = rng.normal(size=(300, 3))
    u0 /= np.linalg.norm(u0, axis=1)[:, None]
    out = []
    for r in ratios:
        u = u0 * r * np.linalg.norm(D)
        e = gr.exact_field(D, m, Q, 1e-4, u) - gr.taylor_field(D, m, Q, 1e-4, u)
        out.append(np.abs(e).max())
    return np.array(out)


def test_monopole_remainder_is_third_order():
    e = _errors(2.0, np.zeros(6), (0.04, 0.02, 0.01))
    rate = e[:-1] / e[1:]
    print("monopole error", e, "halving ratios", rate)
    assert (rate > 7.5).all() and (rate < 8.6).all()


@pytest.mark.parametrize("seed", [2, 3])
def test_quadrupole_remainder_is_at_least_second_order(seed):
    Q = np.random.default_rng(seed).normal(size=6) * 0.3
    e = _errors(2.0, Q, (0.04, 0.02, 0.01), seed)
    rate = e[:-1] / e[1:]
    print("quadrupole error", e, "halving ratios", rate)
    assert (rate > 3.8).all()


def test_expansion_is_exact_at_the_centre():
    Q = np.random.default_rng(4).normal(size=6)
    D = np.array([[1.0, 2.0, -0.5], [-3.0, 0.2, 0.7]])
    m = np.array([1.5, 0.4])
    u = np.zeros((1, 3))
    np.testing.assert_allclose(gr.taylor_field(D, m, np.stack([Q, -Q]), 1e-3, u),
                               gr.exact_field(D, m, np.stack([Q, -Q]), 1e-3, u), rtol=1e-13)
