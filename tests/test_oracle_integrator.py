"""Pins the KDK leapfrog (nbStepSystemPlain, nbody/src/nbody_plain.c:268-364)
through its defining properties: time reversibility, second-order energy
error, exact two-body motion in the limit of small steps."""
import numpy as np

from helpers import plummer_bodies
from oracle import oracle as orc


def test_time_reversible():
    pos, vel, m = plummer_bodies(300, seed=5)
    ctx = orc.make_ctx(eps2=1e-3, theta=0.0, timestep=1e-3)
    o = orc.Oracle(ctx, pos, vel, m)
    o.gravity()
    for _ in range(50):
        o.step()
    back = orc.Oracle(ctx, o.pos, -o.vel, m)
    back.gravity()
    for _ in range(50):
        back.step()
    assert np.abs(back.pos - pos).max() < 1e-10
    assert np.abs(-back.vel - vel).max() < 1e-8


def _energy(o, m, eps2):
    pos, vel = o.pos, o.vel
    k = 0.5 * np.sum(m * np.sum(vel * vel, 1))
    d = pos[:, None, :] - pos[None, :, :]
    r = np.sqrt(np.sum(d * d, 2) + eps2)
    w = -0.5 * np.sum(m[:, None] * m[None, :] / r * (1 - np.eye(len(m))))
    return k + w


def test_energy_error_is_second_order():
    pos, vel, m = plummer_bodies(200, seed=8)
    eps2 = 1e-3
    e0 = None
    errs = []
    for dt, n in ((4e-3, 50), (2e-3, 100)):
        ctx = orc.make_ctx(eps2=eps2, theta=0.0, timestep=dt)
        o = orc.Oracle(ctx, pos, vel, m)
        o.gravity()
        e0 = _energy(o, m, eps2)
        emax = 0.0
        for _ in range(n):
            o.step()
            emax = max(emax, abs(_energy(o, m, eps2) - e0))
        errs.append(emax / abs(e0))
    assert errs[1] < errs[0] / 2.5        # ~4x for a second-order method


def test_circular_orbit_in_point_potential():
    """Light test body on a circular orbit of radius 1 about a heavy body
    (G M = 1): after one period at dt = T/2000 it is back to ~1e-5."""
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    vel = np.array([[0.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    m = np.array([1.0, 1e-12])
    T = 2 * np.pi
    ctx = orc.make_ctx(eps2=0.0, theta=0.0, timestep=T / 2000)
    o = orc.Oracle(ctx, pos, vel, m)
    o.gravity()
    for _ in range(2000):
        o.step()
    assert np.linalg.norm(o.pos[1] - pos[1]) < 5e-5


def test_kick_drift_kick_sequence_for_free_particle():
    """With eps2 huge and tiny masses the force vanishes: x advances by v dt."""
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 3.0]])
    vel = np.array([[0.5, -1.0, 2.0], [0.0, 0.0, 1.0]])
    m = np.array([1e-300, 1e-300])
    ctx = orc.make_ctx(eps2=1.0, theta=0.5, timestep=0.25)
    o = orc.Oracle(ctx, pos, vel, m)
    o.gravity()
    for _ in range(4):
        o.step()
    np.testing.assert_allclose(o.pos, pos + vel, rtol=0, atol=1e-15)
