"""Gradients of a stream likelihood with respect to the dwarf parameters.

The N-body part is the library's hand-written adjoint leapfrog
(nbhip_adjoint_forward / nbhip_adjoint_backward: FP64, tree frozen per
step, checkpoints every `every` steps), exposed to torch as
``KDKRun``, a torch.autograd.Function.  Everything around it -- initial
conditions, orbit, loss -- is written in torch so autograd chains through
it.  Nothing here exists in the reference; it serves a gradient-based
version of the likelihood search that the reference runs by sampling.

Parameters follow the sample workunit's argument vector
(nbody/sample_workunits/for_developers.lua:399-405, 456-480):
    p = (evolveTime T, time_ratio tau, rscale_l a_b, light_r_ratio rr,
         mass_l M_b, light_mass_ratio mr)
    revOrbTime = T / tau, a_d = a_b (1 - rr) / rr, M_d = M_b (1 - mr) / mr
and its two-Plummer mixed dwarf (:323-337).

Project-defined parts (the reference has no differentiable counterpart):
  * initial conditions are a fixed set of unit draws mapped smoothly by the
    parameters: radius by the Plummer inverse CDF truncated at the
    reference's sampling bound 50 (a_b + a_d); speed q * v_esc(r) in the
    total dwarf potential with q drawn once from q^2 (1 - q^2)^(7/2) (the
    Plummer distribution function in its own potential); isotropic
    directions; each component recentred on its own centre of mass
    (nbody/src/nbody_mixeddwarf.c:649-690).  The reference instead samples
    each component's Eddington DF in the total potential by rejection, which
    is not differentiable in the parameters;
  * the step counts are fixed at the reference point: nbody steps
    n = round(T*/dt*), dt = T / n; reverse-orbit steps n_o = round(T*_rev /
    (dt*/10)), step T_rev / n_o (the reference loops "t >= -tstop", whose
    step count is piecewise constant in T);
  * the loss is the squared difference of Gaussian-smoothed (sigma = half a
    bin) normalised lambda histograms of the baryons (50 bins on
    [-150, 150] degrees in the Orphan frame phi, theta, psi =
    128.79, 54.39, 90.70, as the sample workunit's histogram;
    nbXYZToLambdaBeta, nbody/src/nbody_coordinates.c:129-164).
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import models
from . import nbody as nb
from .histogram import HistogramParams

F64 = torch.float64


# ---------------------------------------------------------------------------
# N-body evolution through the C ABI
# ---------------------------------------------------------------------------
@dataclass
class EvolveConfig:
    nsteps: int
    eps2: float
    theta: float = 0.5
    criterion: str = "bh86"
    use_quad: bool = True
    every: int = 50
    potential: Optional[nb.Potential] = None
    device: int = 0

    def ctx(self, dt: float) -> nb.NBodyCtx:
        c = nb.NBodyCtx.create(timestep=dt, timeEvolve=dt * self.nsteps, eps2=self.eps2,
                               theta=self.theta if self.criterion != "exact" else -1.0,
                               criterion=self.criterion, useQuad=self.use_quad)
        # keep dt exactly (NBodyCtx.create corrects it from timeEvolve)
        return dataclasses.replace(c, timestep=dt, nStep=self.nsteps)


def evolve(pos: np.ndarray, vel: np.ndarray, mass: np.ndarray, dt: float, cfg: EvolveConfig):
    """Plain forward run (no checkpoints): final positions and velocities."""
    st = nb.NBodyState(cfg.ctx(dt), cfg.potential, pos, vel, mass, device=cfg.device,
                       prestep=False)
    st.run(cfg.nsteps)
    x, v, _ = st.bodies()
    st.close()
    return x, v


class KDKRun(torch.autograd.Function):
    """(pos [n,3], vel [n,3], mass [n], dt []) -> (pos_N, vel_N) after
    cfg.nsteps KDK steps; backward runs nbhip_adjoint_backward."""

    @staticmethod
    def forward(ctx, pos, vel, mass, dt, cfg: EvolveConfig):
        st = nb.NBodyState(cfg.ctx(float(dt)), cfg.potential, pos.detach().numpy(),
                           vel.detach().numpy(), mass.detach().numpy(), device=cfg.device,
                           prestep=False)
        st.adjoint_forward(cfg.nsteps, cfg.every)
        x, v, _ = st.bodies()
        ctx.st = st
        return torch.from_numpy(x), torch.from_numpy(v)

    @staticmethod
    def backward(ctx, gx, gv):
        st = ctx.st
        gp, gvel, gm, gdt = st.adjoint_backward(gx.detach().contiguous().numpy(),
                                                gv.detach().contiguous().numpy())
        st.close()
        ctx.st = None
        return (torch.from_numpy(gp), torch.from_numpy(gvel), torch.from_numpy(gm),
                torch.tensor(float(gdt[0]), dtype=F64), None)


# ---------------------------------------------------------------------------
# External potential and reverse orbit in torch
# ---------------------------------------------------------------------------
def ext_accel_t(pot: nb.Potential, x: torch.Tensor) -> torch.Tensor:
    """nbExtAcceleration (nbody/src/nbody_potential.c:627-762) on [..., 3]."""
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    ab = torch.sqrt(Z * Z + (Y * Y + X * X))
    r = torch.clamp(ab, min=0.00390625)
    a = torch.zeros_like(x)
    for d in (pot.disk, pot.disk2):
        if d.type == 1:
            M, A, B = d.mass, d.scaleLength, d.scaleHeight
            zp = torch.sqrt(Z * Z + B * B)
            azp = A + zp
            rth = (X * X + Y * Y + azp * azp) ** 1.5
            a = a + torch.stack([-M * X / rth, -M * Y / rth, -M * Z * azp / (zp * rth)], -1)
    h = pot.halo
    if h.type == 1:
        k = -2.0 * h.vhalo ** 2 / (h.scaleLength ** 2 + X * X + Y * Y + (Z / h.flattenZ) ** 2)
        a = a + torch.stack([k * X, k * Y, k * Z / h.flattenZ ** 2], -1)
    elif h.type == 2:
        M = h.vhalo ** 2 * h.scaleLength / 0.2162165954
        c = (-M / (r * r)) * (torch.log((h.scaleLength + r) / h.scaleLength) / r
                              - 1.0 / (h.scaleLength + r))
        a = a + c[..., None] * x
    s = pot.spherical
    if s.type == 1:
        a = a + (-s.mass / (r * (s.scale + r) ** 2))[..., None] * x
    elif s.type == 2:
        a = a + (-s.mass / torch.sqrt(s.scale ** 2 + r * r) ** 3)[..., None] * x
    return a


def reverse_orbit_t(pot: nb.Potential, pos, vel, t_rev: torch.Tensor, nsteps: int):
    """nbReverseOrbit's KDK scheme (nbody/src/nbody_orbit_integrator.c:44-86)
    with a fixed step count and step t_rev / nsteps; pos, vel [..., 3]."""
    dt = t_rev / nsteps
    half = 0.5 * dt
    x = pos.clone()
    v = -vel
    a = ext_accel_t(pot, x)
    for _ in range(nsteps):
        v = v + half[..., None] * a
        x = x + dt[..., None] * v
        a = ext_accel_t(pot, x)
        v = v + half[..., None] * a
    return x, -v


class ReverseOrbit(torch.autograd.Function):
    """reverse_orbit_t through the library: the end point and its derivative
    with respect to t_rev come from one forward-mode pass in native code
    (nbhip_orbit_tangent) instead of a torch graph of every step."""

    @staticmethod
    def forward(ctx, t_rev, pos, vel, pot: nb.Potential, nsteps: int):
        x, v, dx, dv = nb.orbit_tangent(pot, pos.detach().numpy(), vel.detach().numpy(),
                                        float(t_rev), nsteps)
        ctx.save_for_backward(torch.from_numpy(dx), torch.from_numpy(dv))
        return torch.from_numpy(x), torch.from_numpy(v)

    @staticmethod
    def backward(ctx, gx, gv):
        dx, dv = ctx.saved_tensors
        return (gx * dx).sum() + (gv * dv).sum(), None, None, None, None


# ---------------------------------------------------------------------------
# Differentiable two-Plummer dwarf
# ---------------------------------------------------------------------------
def _q_table(nq: int = 20001):
    """Inverse CDF of q on [0, 1) with density proportional to q^2 (1-q^2)^3.5."""
    q = np.linspace(0.0, 1.0, nq)
    pdf = q * q * (1.0 - q * q) ** 3.5
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(q))])
    return cdf / cdf[-1], q


@dataclass
class DwarfDraws:
    """Fixed unit draws of one dwarf: baryons first, then dark matter."""
    n: int
    n_b: int
    u_r: np.ndarray        # radius quantiles in (0, 1)
    dir_x: np.ndarray      # unit position directions [n, 3]
    q: np.ndarray          # speed / escape speed
    dir_v: np.ndarray      # unit velocity directions [n, 3]

    @staticmethod
    def create(n: int, n_b: int, seed: int) -> "DwarfDraws":
        rng = np.random.default_rng(seed)

        def dirs(k):
            v = rng.normal(size=(k, 3))
            return v / np.linalg.norm(v, axis=1, keepdims=True)

        cdf, qg = _q_table()
        q = np.interp(rng.uniform(size=n), cdf, qg)
        return DwarfDraws(n, n_b, rng.uniform(1e-6, 1.0 - 1e-6, size=n), dirs(n), q, dirs(n))


def plummer_radius(u, a, rmax):
    """Plummer inverse CDF, truncated: r with M(<r)/M(<rmax) = u."""
    y = rmax / a
    fmax = y ** 3 / (1.0 + y * y) ** 1.5
    f = u * fmax
    return a / torch.sqrt(f ** (-2.0 / 3.0) - 1.0)


def dwarf_parameters(p: torch.Tensor):
    """(T, T_rev, a_b, a_d, M_b, M_d) from the workunit parameter vector."""
    T, tau, a_b, rr, M_b, mr = p.unbind()
    return T, T / tau, a_b, a_b * (1.0 - rr) / rr, M_b, M_b * (1.0 - mr) / mr


def dwarf_ic(p: torch.Tensor, draws: DwarfDraws, centre_pos, centre_vel):
    """Positions, velocities and masses of the dwarf (centred at the given
    phase-space point) as smooth functions of p."""
    _, _, a_b, a_d, M_b, M_d = dwarf_parameters(p)
    nb_, nd = draws.n_b, draws.n - draws.n_b
    rmax = 50.0 * (a_b + a_d)
    u = torch.from_numpy(draws.u_r)
    r = torch.cat([plummer_radius(u[:nb_], a_b, rmax), plummer_radius(u[nb_:], a_d, rmax)])
    psi = M_b / torch.sqrt(r * r + a_b * a_b) + M_d / torch.sqrt(r * r + a_d * a_d)
    speed = torch.from_numpy(draws.q) * torch.sqrt(2.0 * psi)
    x = r[:, None] * torch.from_numpy(draws.dir_x)
    v = speed[:, None] * torch.from_numpy(draws.dir_v)
    parts_x, parts_v = [], []
    for sl in (slice(0, nb_), slice(nb_, draws.n)):
        parts_x.append(x[sl] - x[sl].mean(0))
        parts_v.append(v[sl] - v[sl].mean(0))
    x = torch.cat(parts_x) + centre_pos
    v = torch.cat(parts_v) + centre_vel
    m = torch.cat([(M_b / nb_).expand(nb_), (M_d / nd).expand(nd)])
    return x, v, m


# ---------------------------------------------------------------------------
# Smoothed lambda histogram
# ---------------------------------------------------------------------------
def lambda_deg(x: torch.Tensor, spec: HistogramParams) -> torch.Tensor:
    """nbXYZToLambdaBeta's lambda (nbody/src/nbody_coordinates.c:129-164)."""
    (r11, r12, r13), (r21, r22, r23) = spec.rotation()[:2].tolist()
    X = x[:, 0] + spec.sunGCDist
    tx = r11 * X + r12 * x[:, 1] + r13 * x[:, 2]
    ty = r21 * X + r22 * x[:, 1] + r23 * x[:, 2]
    return torch.rad2deg(torch.atan2(ty, tx))


def smoothed_histogram(lam: torch.Tensor, spec: HistogramParams, sigma_bins: float = 0.5):
    """Each body spreads a Gaussian (sigma = sigma_bins bin widths) over the
    lambda bins; normalised by the body count."""
    w = (spec.lambdaEnd - spec.lambdaStart) / spec.lambdaBins
    edges = spec.lambdaStart + w * torch.arange(spec.lambdaBins + 1, dtype=F64)
    z = (edges[None, :] - lam[:, None]) / (sigma_bins * w)
    cdf = torch.special.ndtr(z)
    return (cdf[:, 1:] - cdf[:, :-1]).sum(0) / lam.shape[0]


# ---------------------------------------------------------------------------
# The whole model
# ---------------------------------------------------------------------------
@dataclass
class StreamModel:
    """Likelihood-search forward model: parameters -> smoothed histogram."""
    draws: DwarfDraws
    cfg: EvolveConfig
    potential: nb.Potential
    p_ref: np.ndarray                  # fixes the step counts
    dt_ref: float = 1e-3
    orbit_pos: tuple = (218.0, 53.5, 28.6)      # present-day (l, b, r)
    orbit_vel: tuple = (-156.0, 79.0, 107.0)
    spec: HistogramParams = dataclasses.field(default_factory=HistogramParams)

    @property
    def orbit_steps(self) -> int:
        return int(round(self.p_ref[0] / self.p_ref[1] / (self.dt_ref / 10.0)))

    def initial_state(self, p: torch.Tensor):
        T, t_rev = dwarf_parameters(p)[:2]
        x0 = torch.as_tensor(models.lbr_to_cartesian(list(self.orbit_pos)), dtype=F64)
        v0 = torch.tensor(self.orbit_vel, dtype=F64)
        cx, cv = ReverseOrbit.apply(t_rev, x0, v0, self.potential, self.orbit_steps)
        x, v, m = dwarf_ic(p, self.draws, cx, cv)
        return x, v, m, T / self.cfg.nsteps

    def histogram(self, p: torch.Tensor, grad: bool = True) -> torch.Tensor:
        x, v, m, dt = self.initial_state(p)
        if grad:
            xN, _ = KDKRun.apply(x, v, m, dt, self.cfg)
        else:
            xN = torch.from_numpy(evolve(x.detach().numpy(), v.detach().numpy(),
                                         m.detach().numpy(), float(dt), self.cfg)[0])
        return smoothed_histogram(lambda_deg(xN[:self.draws.n_b], self.spec), self.spec)

    def loss(self, p: torch.Tensor, target: torch.Tensor, grad: bool = True) -> torch.Tensor:
        return ((self.histogram(p, grad) - target) ** 2).sum()

    def value_and_grad(self, p, target):
        pt = torch.tensor(np.asarray(p, dtype=np.float64), requires_grad=True)
        L = self.loss(pt, target, grad=True)
        L.backward()
        return float(L.detach()), pt.grad.numpy().copy()


def make_model(n: int = 20000, p_ref=(1.0, 1.0, 0.2, 0.2, 12.0, 0.2), seed: int = 34086709,
               criterion: str = "bh86", every: int = 50, dt: float = 1e-3,
               theta: float = 0.5) -> StreamModel:
    """The differentiation workload: two-Plummer dwarf of n bodies (half
    baryons), T/dt KDK steps in the sample workunit's Milky Way, eps2 from
    calculateEps2 at the reference parameters."""
    p_ref = np.asarray(p_ref, dtype=np.float64)
    _, _, a_b, a_d, M_b, M_d = (float(t) for t in dwarf_parameters(torch.from_numpy(p_ref)))
    eps2 = models.calculate_eps2(n, a_b, a_d, M_b, M_d)
    pot = nb.milky_way_potential()
    cfg = EvolveConfig(nsteps=int(round(p_ref[0] / dt)), eps2=eps2, theta=theta,
                       criterion=criterion, use_quad=criterion != "exact", every=every,
                       potential=pot)
    return StreamModel(DwarfDraws.create(n, n // 2, seed), cfg, pot, p_ref, dt)
