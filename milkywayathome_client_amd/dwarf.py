"""Two-component dwarf initial conditions (Plummer baryons + NFW or Plummer
dark matter) for the two-component workloads.

The model is the reference's ``predefinedModels.mixeddwarf``
(nbody/src/nbody_mixeddwarf.c:765-1026): the same component densities and
potentials (nbody/src/nbody_dwarf_potential.c), the same NFW normalisation
(M200 input, mass reset to M(<5 r200); set_vars / get_extra_nfw_mass,
:696-762), the same sampling bounds, isotropic velocities from each
component's Eddington distribution function in the total potential, and the
per-component centre-of-mass correction (:649-690).

The sampler itself is this project's own: radii by inverse-CDF, the
distribution functions tabulated once by Eddington inversion, and speeds by
vectorised rejection against that table.  The reference draws every body by
per-trial numerical quadrature, which takes hours at 10^6 bodies; the
outputs are therefore statistically equivalent, not stream-identical.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .models import DSFMT

PCRIT = 0.000679087369829744220469326744094105320596648627735869652
VOL_PCRIT = 0.568910904587397184785763397846734505212216314432372653620


@dataclass
class Component:
    kind: str          # "plummer" or "nfw"
    mass: float        # Plummer: total mass; NFW: M200 on input
    scale: float
    p0: float = 0.0
    r200: float = 0.0

    def setup(self, other_scale: float) -> float:
        """set_vars + sampling bound (+ NFW mass reset); returns the bound."""
        r200 = np.cbrt(self.mass / VOL_PCRIT)
        c = r200 / self.scale
        term = np.log(1.0 + c) - c / (1.0 + c)
        self.p0 = 200.0 * c ** 3 * PCRIT / (3.0 * term)
        self.r200 = r200
        if self.kind == "plummer":
            return 50.0 * (self.scale + other_scale)
        bound = 5.0 * r200
        a = self.scale
        self.mass = 4.0 * np.pi * self.p0 * a ** 3 * (np.log((a + bound) / a) - bound / (a + bound))
        return bound

    def density(self, r):
        a = self.scale
        if self.kind == "plummer":
            return (3.0 / (4.0 * np.pi)) * (self.mass / a ** 3) * (1.0 + (r / a) ** 2) ** -2.5
        R = r / a
        return self.p0 / (R * (1.0 + R) ** 2)

    def ddensity_dr(self, r):
        a = self.scale
        if self.kind == "plummer":
            return (3.0 / (4.0 * np.pi)) * (self.mass / a ** 3) * (-5.0 * r / a ** 2) * \
                (1.0 + (r / a) ** 2) ** -3.5
        R = r / a
        return -self.p0 / a * (1.0 + 3.0 * R) / (R ** 2 * (1.0 + R) ** 3)

    def psi(self, r):
        a = self.scale
        if self.kind == "plummer":
            return self.mass / np.sqrt(r * r + a * a)
        R = r / a
        return 4.0 * np.pi * a * a * self.p0 * np.log1p(R) / R

    def dpsi_dr(self, r):
        a = self.scale
        if self.kind == "plummer":
            return -self.mass * r / (r * r + a * a) ** 1.5
        R = r / a
        return 4.0 * np.pi * a * self.p0 * (R / (1.0 + R) - np.log1p(R)) / R ** 2

    def enclosed_fraction(self, r, bound):
        a = self.scale
        if self.kind == "plummer":
            f = lambda x: x ** 3 / (x * x + a * a) ** 1.5
        else:
            f = lambda x: np.log1p(x / a) - x / (a + x)
        return f(r) / f(bound)


def _eddington_table(target: Component, comps, rgrid):
    """log E grid and f(E) of `target` in the potential of all comps."""
    psi = sum(c.psi(rgrid) for c in comps)                 # decreasing in r
    dpsi = sum(c.dpsi_dr(rgrid) for c in comps)
    g = target.ddensity_dr(rgrid) / dpsi                    # d rho / d psi >= 0
    g = np.maximum(g, 0.0)
    # ascending psi for interpolation
    ps = psi[::-1]
    gs = g[::-1]
    E = ps[1:]                                              # energies on the grid
    # F(E) = 2 int_0^sqrt(E) g(E - t^2) dt  (Gauss-Legendre in t)
    xg, wg = np.polynomial.legendre.leggauss(96)
    t = 0.5 * (xg + 1.0)                                    # in (0,1)
    sq = np.sqrt(E)
    F = np.zeros_like(E)
    for tk, wk in zip(t, wg):
        arg = E - (tk * sq) ** 2
        F += wk * 0.5 * sq * np.interp(arg, ps, gs)
    F *= 2.0
    # far out psi stops changing in FP64: keep strictly increasing energies
    keep = np.concatenate([[True], np.diff(E) > 0.0])
    E, F = E[keep], F[keep]
    dF = np.gradient(F, E)
    f = np.maximum(dF, 0.0) / (np.sqrt(8.0) * np.pi ** 2)
    return E, f


def _unit_vectors(prng: DSFMT, n: int) -> np.ndarray:
    """Isotropic directions by rejection from the cube (get_components,
    nbody/src/nbody_mixeddwarf.c:620-647)."""
    out = np.empty((n, 3))
    filled = 0
    while filled < n:
        m = int((n - filled) * 2.1) + 16
        v = -1.0 + 2.0 * prng.draw(3 * m, "open_open").reshape(m, 3)
        r2 = np.einsum("ij,ij->i", v, v)
        v = v[(r2 <= 1.0) & (r2 > 0.0)]
        v /= np.sqrt(np.einsum("ij,ij->i", v, v))[:, None]
        k = min(v.shape[0], n - filled)
        out[filled:filled + k] = v[:k]
        filled += k
    return out


def mixed_dwarf(nbody: int, nbody_baryon: int, light: Component, dark: Component, seed: int,
                position=(0.0, 0.0, 0.0), velocity=(0.0, 0.0, 0.0)):
    """Returns (pos (n,3), vel (n,3), mass (n,), type (n,)); baryons first
    (type 1), dark matter after (type -1, ignored by the histogram)."""
    prng = DSFMT(seed)
    bound_l = light.setup(dark.scale)
    bound_d = dark.setup(light.scale)
    comps = [light, dark]
    n_l, n_d = nbody_baryon, nbody - nbody_baryon
    amin = min(light.scale, dark.scale)
    rgrid = np.geomspace(1e-7 * amin, 1e5 * max(bound_l, bound_d), 6000)

    radii, speeds, masses = [], [], []
    for comp, n, bound in ((light, n_l, bound_l), (dark, n_d, bound_d)):
        if n == 0:
            continue
        # radii: inverse of the truncated enclosed-mass fraction
        rr = np.geomspace(1e-9 * comp.scale, bound, 20000)
        cdf = comp.enclosed_fraction(rr, bound)
        u = prng.draw(n, "open_open")
        r = np.interp(u, cdf, rr)
        # speeds: rejection on v^2 f(psi(r) - v^2/2)
        E, f = _eddington_table(comp, comps, rgrid)
        logE, logf = np.log(E), np.log(np.maximum(f, 1e-300))
        fE = lambda e: np.exp(np.interp(np.log(np.maximum(e, E[0])), logE, logf))
        psi_r = sum(c.psi(r) for c in comps)
        vesc = 0.99 * np.sqrt(2.0 * psi_r)
        vg = np.linspace(0.0, 1.0, 65)[1:]
        gmax = np.zeros(n)
        for x in vg:
            v = x * vesc
            gmax = np.maximum(gmax, v * v * fE(psi_r - 0.5 * v * v))
        gmax *= 1.3
        v = np.zeros(n)
        todo = np.arange(n)
        for _ in range(200):
            if todo.size == 0:
                break
            u1 = prng.draw(todo.size, "open_open")
            u2 = prng.draw(todo.size, "open_open")
            vt = u1 * vesc[todo]
            val = vt * vt * fE(psi_r[todo] - 0.5 * vt * vt)
            ok = val > u2 * gmax[todo]
            v[todo[ok]] = vt[ok]
            todo = todo[~ok]
        if todo.size:
            raise RuntimeError(f"dwarf sampler: {todo.size} speeds not accepted")
        radii.append(r)
        speeds.append(v)
        masses.append(np.full(n, comp.mass / n))

    r = np.concatenate(radii)
    v = np.concatenate(speeds)
    m = np.concatenate(masses)
    pos = _unit_vectors(prng, nbody) * r[:, None]
    vel = _unit_vectors(prng, nbody) * v[:, None]
    # per-component centre of mass and momentum to the orbit point
    rs = np.asarray(position, dtype=np.float64)
    vs = np.asarray(velocity, dtype=np.float64)
    for lo, hi, comp in ((0, n_l, light), (n_l, nbody, dark)):
        if hi <= lo:
            continue
        mm = m[lo:hi, None]
        pos[lo:hi] += rs - (mm * pos[lo:hi]).sum(0) / comp.mass
        vel[lo:hi] += vs - (mm * vel[lo:hi]).sum(0) / comp.mass
    typ = np.where(np.arange(nbody) < n_l, 1, -1).astype(np.int32)
    return pos, vel, m, typ


def two_component_dwarf(nbody: int, seed: int = 34086709, mass_l: float = 12.0,
                        rscale_l: float = 0.2, mass_d: float = 60.0, rscale_d: float = 0.8,
                        dark: str = "nfw", baryon_fraction: float = 0.5,
                        position=(0.0, 0.0, 0.0), velocity=(0.0, 0.0, 0.0)):
    nb = int(round(nbody * baryon_fraction))
    return mixed_dwarf(nbody, nb, Component("plummer", mass_l, rscale_l),
                       Component(dark, mass_d, rscale_d), seed, position, velocity)
