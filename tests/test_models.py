"""Initial-condition layer: Plummer sampler (nbody/src/nbody_plummer.c),
two-component dwarf (statistics of nbody/tests/mixeddwarf_test.c), reverse
orbit (nbody/src/nbody_orbit_integrator.c) and softening helpers."""
import math

import numpy as np
import pytest

from milkywayathome_client_amd import dwarf, models, nbody as nb
from milkywayathome_client_amd.models import DSFMT, plummer
from oracle import oracle as orc


def test_plummer_is_deterministic_and_equal_mass():
    a = plummer(1000, DSFMT(42), 12.0, 0.2)
    b = plummer(1000, DSFMT(42), 12.0, 0.2)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert np.all(a[2] == 12.0 / 1000) and np.all(a[3] == 1)
    assert np.all(plummer(10, DSFMT(1), 1.0, 1.0, ignore=True)[3] == -1)


def test_plummer_radii_follow_the_plummer_mass_profile():
    """Radii are r = a / sqrt(u^(-2/3) - 1): the enclosed fraction
    M(<r)/M = r^3 / (r^2 + a^2)^(3/2) is uniform; KS distance small."""
    a = 0.2
    pos, vel, m, _ = plummer(20000, DSFMT(7), 12.0, a)
    r = np.sort(np.linalg.norm(pos, axis=1))
    F = r ** 3 / (r * r + a * a) ** 1.5
    ks = np.max(np.abs(F - (np.arange(1, r.size + 1) - 0.5) / r.size))
    assert ks < 1.63 / math.sqrt(r.size)   # 1% KS level


def test_plummer_directions_isotropic_and_shifted():
    pos, vel, _, _ = plummer(20000, DSFMT(9), 12.0, 0.2, (1.0, 2.0, 3.0), (-5.0, 0.0, 5.0))
    assert np.abs(np.median(pos, 0) - [1.0, 2.0, 3.0]).max() < 0.01
    u = (pos - [1.0, 2.0, 3.0]) / np.linalg.norm(pos - [1.0, 2.0, 3.0], axis=1)[:, None]
    assert np.abs(u.mean(0)).max() < 0.03
    assert np.abs(np.median(vel, 0) - [-5.0, 0.0, 5.0]).max() < 0.2


def test_plummer_speed_bound_and_virial():
    """Velocities are drawn below escape speed of the sampler's potential
    sqrt(2 M / a) (1 + r^2)^(-1/4) in units of the scale radius."""
    M, a = 12.0, 0.2
    pos, vel, m, _ = plummer(4000, DSFMT(3), M, a)
    r = np.linalg.norm(pos, axis=1) / a
    v = np.linalg.norm(vel, axis=1)
    vesc = math.sqrt(2.0 * M / a) * (1.0 + r * r) ** -0.25
    assert np.all(v <= vesc * (1 + 1e-12))


def test_dwarf_virial_and_component_centres():
    """mixeddwarf_test.c:583-700: virial ratio within 5 %, each component's
    centre of mass and momentum on the orbit point to 1e-6."""
    rs, vs = np.array([-35.0, -28.0, 25.0]), np.array([50.0, 130.0, -10.0])
    pos, vel, m, typ = dwarf.two_component_dwarf(20000, seed=11, position=rs, velocity=vs)
    assert np.all(typ[:10000] == 1) and np.all(typ[10000:] == -1)
    for sl in (slice(0, 10000), slice(10000, 20000)):
        mm = m[sl, None]
        np.testing.assert_allclose((mm * pos[sl]).sum(0) / mm.sum(), rs, atol=1e-6)
        np.testing.assert_allclose((mm * vel[sl]).sum(0) / mm.sum(), vs, atol=1e-6)
    d, u = pos - rs, vel - vs
    K = 0.5 * np.sum(m * np.sum(u * u, 1))
    idx = np.arange(0, 20000, 5)
    a = orc.direct_accel(d, m, 1e-8, idx)
    W = 5.0 * np.sum(m[idx] * np.sum(d[idx] * a, 1))
    assert abs(2 * K / abs(W) - 1.0) < 0.05


def test_nfw_mass_is_reset_to_five_r200():
    c = dwarf.Component("nfw", 60.0, 0.8)
    bound = c.setup(0.2)
    assert math.isclose(bound, 5.0 * np.cbrt(60.0 / dwarf.VOL_PCRIT))
    assert c.mass > 60.0


def test_reverse_orbit_round_trip():
    """Integrating the reversed endpoint forward with the same leapfrog
    returns to the start: the reverse orbit is a time-reversed KDK."""
    pot = nb.milky_way_potential()
    x0 = models.lbr_to_cartesian([218.0, 53.5, 28.6])
    v0 = np.array([-156.0, 79.0, 107.0])
    fx, fv = models.reverse_orbit(pot, x0, v0, 1.0, 1e-4)
    # forward: reverse_orbit of (fx, -fv) backwards in time is a forward run
    bx, bv = models.reverse_orbit(pot, fx, -fv, 1.0, 1e-4)
    assert np.linalg.norm(bx - x0) < 1e-6 * np.linalg.norm(x0)
    assert np.linalg.norm(-bv - v0) < 1e-6 * np.linalg.norm(v0)


def test_lbr_to_cartesian_closed_form():
    x = models.lbr_to_cartesian([90.0, 0.0, 10.0], 8.0)
    np.testing.assert_allclose(x, [-8.0, 10.0, 0.0], atol=1e-12)
    x = models.lbr_to_cartesian([0.0, 90.0, 5.0], 8.0)
    np.testing.assert_allclose(x, [-8.0, 0.0, 5.0], atol=1e-12)


def test_softening_and_timestep_helpers():
    # single Plummer sphere of scale a: the Henon length unit is 16 a / (3 pi)
    rv = models.load_models().nbm_virial_radius(0.2, 1.0, 1.0, 0.0)
    assert math.isclose(rv, 16 * 0.2 / (3 * math.pi), rel_tol=1e-6)
    eps2 = models.calculate_eps2(10000, 0.2, 1.0, 12.0, 0.0)
    assert math.isclose(eps2, (rv * 0.98 * 10000 ** -0.26) ** 2, rel_tol=1e-12)
    dt = models.calculate_timestep(16.0, 0.2)
    assert math.isclose(dt, 0.01 * math.sqrt(4.0 / 3.0 * math.pi * 0.008 / 16.0))


def test_virial_radius_of_one_plummer():
    """nbody/tests/virial_test.c: with one component massless the virial
    radius is the single Plummer sphere's, 16 a / (3 pi), to 1e-5 (200 random
    scale radii in [0.1, 5] and masses in [1, 20])."""
    rng = np.random.default_rng(17)
    for _ in range(200):
        a_b, a_d = rng.uniform(0.1, 5.0, 2)
        m_b, m_d = rng.uniform(1.0, 20.0, 2)
        if rng.uniform() < 0.5:
            m_b, lhs = 0.0, 16.0 * a_d / 3.0 / np.pi
        else:
            m_d, lhs = 0.0, 16.0 * a_b / 3.0 / np.pi
        assert abs(models.virial_radius(a_b, a_d, m_b, m_d) - lhs) <= 1e-5
