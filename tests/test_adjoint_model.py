"""CPU checks of the differentiable stream model around the HIP adjoint
(milkywayathome_client_amd/adjoint.py): the torch reverse orbit against the
C restatement of nbReverseOrbit, the lambda frame against its defining
property, the smoothed histogram's limits, and torch's gradients of the
initial conditions against finite differences (gradcheck)."""
import math

import numpy as np
import pytest
import torch

from milkywayathome_client_amd import adjoint as adj
from milkywayathome_client_amd import models
from milkywayathome_client_amd import nbody as nb

F64 = torch.float64


def test_torch_orbit_equals_c_reverse_orbit():
    pot = nb.milky_way_potential()
    x0 = models.lbr_to_cartesian([218.0, 53.5, 28.6])
    v0 = np.array([-156.0, 79.0, 107.0])
    fx, fv = models.reverse_orbit(pot, x0, v0, 0.2, 1e-4)     # loop runs t = 0 .. -0.2
    tx, tv = adj.reverse_orbit_t(pot, torch.tensor(x0, dtype=F64), torch.tensor(v0, dtype=F64),
                                 torch.tensor(2001 * 1e-4, dtype=F64), 2001)
    np.testing.assert_allclose(tx.numpy(), fx, rtol=1e-12)
    np.testing.assert_allclose(tv.numpy(), fv, rtol=1e-12)


def test_native_orbit_equals_c_reverse_orbit():
    """nbhip_orbit_tangent's end point (host code in the library, no device)
    against the C restatement of nbReverseOrbit over the same steps."""
    pot = nb.milky_way_potential()
    x0 = models.lbr_to_cartesian([218.0, 53.5, 28.6])
    v0 = np.array([-156.0, 79.0, 107.0])
    fx, fv = models.reverse_orbit(pot, x0, v0, 0.2, 1e-4)     # loop runs t = 0 .. -0.2
    nx, nv, _, _ = nb.orbit_tangent(pot, x0, v0, 2001 * 1e-4, 2001)
    np.testing.assert_allclose(nx, fx, rtol=1e-12)
    np.testing.assert_allclose(nv, fv, rtol=1e-12)


@pytest.mark.parametrize("halo", ["log", "nfw"])
def test_orbit_tangent_matches_autograd_and_differences(halo):
    """d(end point)/d(t_rev) from the forward-mode pass against torch autograd
    through every step of reverse_orbit_t, and against central differences of
    the end point itself."""
    pot = nb.milky_way_potential()
    if halo == "nfw":
        pot = nb.Potential.create(nb.Spherical.plummer(3e4, 0.9),
                                  nb.Disk.miyamotoNagai(2e5, 4.0, 0.5),
                                  nb.Disk.miyamotoNagai(5e4, 2.0, 0.1), nb.Halo.nfw(120.0, 15.0))
    x0 = models.lbr_to_cartesian([218.0, 53.5, 28.6])
    v0 = np.array([-156.0, 79.0, 107.0])
    t_rev, nsteps = 0.7, 700
    nx, nv, dx, dv = nb.orbit_tangent(pot, x0, v0, t_rev, nsteps)
    t = torch.tensor(t_rev, dtype=F64, requires_grad=True)
    tx, tv = adj.reverse_orbit_t(pot, torch.tensor(x0, dtype=F64), torch.tensor(v0, dtype=F64),
                                 t, nsteps)
    np.testing.assert_allclose(tx.detach().numpy(), nx, rtol=1e-12)
    np.testing.assert_allclose(tv.detach().numpy(), nv, rtol=1e-12)
    jac = np.array([torch.autograd.grad(c, t, retain_graph=True)[0].item()
                    for c in list(tx) + list(tv)])
    np.testing.assert_allclose(np.concatenate([dx, dv]), jac, rtol=1e-9,
                               atol=1e-9 * np.abs(jac).max())
    h = 1e-6 * t_rev
    a = nb.orbit_tangent(pot, x0, v0, t_rev + h, nsteps)
    b = nb.orbit_tangent(pot, x0, v0, t_rev - h, nsteps)
    fd = np.concatenate([(a[0] - b[0]) / (2 * h), (a[1] - b[1]) / (2 * h)])
    np.testing.assert_allclose(np.concatenate([dx, dv]), fd, rtol=1e-6,
                               atol=1e-6 * np.abs(fd).max())


def test_orbit_tangent_rejects_bad_arguments():
    with pytest.raises(nb.NBodyError):
        nb.orbit_tangent(nb.milky_way_potential(), np.zeros(3), np.zeros(3), 1.0, 0)


def test_torch_ext_accel_equals_c():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(300, 3)) * 25.0
    for pot in (nb.milky_way_potential(),
                nb.Potential.create(nb.Spherical.plummer(3e4, 0.9),
                                    nb.Disk.miyamotoNagai(2e5, 4.0, 0.5),
                                    nb.Disk.miyamotoNagai(5e4, 2.0, 0.1), nb.Halo.nfw(120.0, 15.0))):
        a = adj.ext_accel_t(pot, torch.tensor(x, dtype=F64)).numpy()
        ref = models.ext_accel(pot, x)
        assert (np.abs(a - ref).max(1) / np.abs(ref).max(1)).max() < 1e-14


def test_lambda_frame():
    spec = adj.HistogramParams()
    R = spec.rotation()
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-15)      # orthonormal
    lam0 = np.array([-120.0, -30.0, 0.0, 45.0, 170.0])
    pts = []
    for l in np.radians(lam0):
        helio = 20.0 * (math.cos(l) * R[0] + math.sin(l) * R[1]) + 3.0 * R[2]
        pts.append(helio - np.array([spec.sunGCDist, 0.0, 0.0]))
    got = adj.lambda_deg(torch.tensor(np.array(pts), dtype=F64), spec).numpy()
    np.testing.assert_allclose(got, lam0, atol=1e-12)


def test_smoothed_histogram_limits():
    spec = adj.HistogramParams()
    rng = np.random.default_rng(2)
    lam = rng.uniform(-140.0, 140.0, size=5000)
    w = (spec.lambdaEnd - spec.lambdaStart) / spec.lambdaBins
    lt = torch.tensor(lam, dtype=F64)
    sharp = adj.smoothed_histogram(lt, spec, sigma_bins=1e-9).numpy()
    plain = np.histogram(lam, bins=spec.lambdaBins, range=(spec.lambdaStart, spec.lambdaEnd))[0]
    np.testing.assert_allclose(sharp, plain / lam.size, atol=1e-12)
    smooth = adj.smoothed_histogram(lt, spec).numpy()
    inside = np.mean([0.5 * (math.erf((spec.lambdaEnd - l) / (0.5 * w * math.sqrt(2)))
                             - math.erf((spec.lambdaStart - l) / (0.5 * w * math.sqrt(2))))
                      for l in lam])
    assert abs(smooth.sum() - inside) < 1e-12


def test_initial_state_gradients():
    model = adj.make_model(n=24, p_ref=(0.02, 1.0, 0.2, 0.2, 12.0, 0.2))
    assert model.orbit_steps == 200

    def f(p):
        x, v, m, dt = model.initial_state(p)
        return x.sum(0), (v * v).sum(0), m[:3], dt

    p = torch.tensor([0.021, 1.1, 0.22, 0.25, 11.0, 0.3], dtype=F64, requires_grad=True)
    assert torch.autograd.gradcheck(f, (p,), eps=1e-6, atol=1e-6, rtol=1e-5)


def test_dwarf_ic_shape_and_scales():
    model = adj.make_model(n=4000)
    p = torch.tensor(model.p_ref)
    x, v, m, dt = model.initial_state(p)
    nb_ = model.draws.n_b
    assert abs(float(dt) - 1e-3) < 1e-18 and model.cfg.nsteps == 1000
    assert abs(float(m[:nb_].sum()) - 12.0) < 1e-12 and abs(float(m[nb_:].sum()) - 48.0) < 1e-12
    xc = x.numpy()
    rb = np.linalg.norm(xc[:nb_] - xc[:nb_].mean(0), axis=1)
    # Plummer half-mass radius 1.305 a (a_b = 0.2)
    assert abs(np.median(rb) / 0.2 - 1.305) < 0.06
    # roughly virialised: 2K/|W| within 25% of 1 (q is the one-component
    # Plummer speed law, not the two-component Eddington DF)
    xs, vs, ms = xc - xc.mean(0), v.numpy() - v.numpy().mean(0), m.numpy()
    K = 0.5 * (ms * (vs ** 2).sum(1)).sum()
    d = np.sqrt(((xs[:, None, :] - xs[None, :, :]) ** 2).sum(2) + 1e-12)
    np.fill_diagonal(d, np.inf)
    W = -0.5 * (ms[:, None] * ms[None, :] / d).sum()
    print("2K/|W|", 2 * K / abs(W))
    assert 0.75 < 2 * K / abs(W) < 1.25
