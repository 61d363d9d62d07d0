"""nbRunSystemPlain's best-likelihood run (nbody/src/nbody_plain.c:75-245,
366-461): one force evaluation, then KDK steps up to nStep; from the step
fraction BestLikeStart on, every step's histogram is compared with the data
histogram and the best (smallest |likelihood|) is kept, with the step time,
its components, the number of improvements, its histogram and, on request,
the bodies.

The histograms come from the GPU state (raw counts from nbhip_histogram,
histograms 1-7 from nbhip_histogram_dispersion); the likelihood is the host
nbSystemLikelihood of likelihood.py.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

from . import likelihood as L

DEFAULT_BEST_CASE = 1e-9                 # nbody/include/nbody_defaults.h:81
DEFAULT_BEST_LIKELIHOOD_START = 0.95     # :67


@dataclass
class BestLikelihood:
    """The reference's st->bestLikelihood* fields: the best likelihood and
    its components (0 for a term not in use), the model time it occurred at
    and how many times it improved; plus the histogram of that step and,
    with keep_bodies, its (pos, vel, acc)."""
    likelihood: float = L.DEFAULT_WORST_CASE
    emd: float = math.nan
    mass: float = math.nan
    beta: float = 0.0
    vel: float = 0.0
    beta_avg: float = 0.0
    vel_avg: float = 0.0
    dist: float = 0.0
    pm_dec: float = 0.0
    pm_ra: float = 0.0
    time: float = 0.0
    step: int = -1
    count: int = 0
    histogram: Optional[L.Histogram] = None
    bodies: Optional[tuple] = None
    history: list = field(default_factory=list)


def clamp_likelihood(like: float) -> float:
    """get_likelihood's clamp (nbody/src/nbody_plain.c:155-163): beyond
    +-DEFAULT_WORST_CASE or NaN -> the worst case, exactly 0 -> the best
    case (1e-9)."""
    if like > L.DEFAULT_WORST_CASE or like < -L.DEFAULT_WORST_CASE or math.isnan(like):
        return L.DEFAULT_WORST_CASE
    if like == 0.0:
        return DEFAULT_BEST_CASE
    return like


def simulated_histogram(st, hp, total_simulated: int, mass_per_particle: float,
                        sim: int = 0, dispersions: bool = True, **disp) -> L.Histogram:
    """nbCreateHistogram for simulation `sim` of a GPU state: the normalised
    counts and, with dispersions, histograms 1-7 (keyword arguments go to
    histogram_dispersion: iter_max, the sigmas and corrections, the
    proper-motion frame)."""
    counts, _ = st.histogram(hp)
    d = st.histogram_dispersion(hp, sim=sim, **disp) if dispersions else None
    return L.from_counts(counts[sim], hp, total_simulated, mass_per_particle, dispersions=d)


def run_best_likelihood(st, data: L.Histogram, hp, total_simulated: int,
                        mass_per_particle: float, best_like_start: float =
                        DEFAULT_BEST_LIKELIHOOD_START, use_beta_disp: bool = False,
                        use_vel_disp: bool = False, use_vlos: bool = False,
                        use_beta_comp: bool = False, use_dist: bool = False,
                        use_prop_mot: bool = False, keep_bodies: bool = False,
                        record: bool = False, **disp) -> BestLikelihood:
    """Run `st` (an NBodyState, one simulation) from its current step to
    ctx.nStep as nbRunSystemPlain does with useBestLike: the force at the
    current positions, then one KDK step at a time; after each step whose
    fraction step / nStep reaches best_like_start, the likelihood of the
    current histogram against `data` (nbSystemLikelihood over st.n bodies,
    clamped as get_likelihood clamps it) replaces the best when its absolute
    value is smaller.  total_simulated and mass_per_particle are
    nbCreateHistogram's: the number of light (not ignored) bodies and the
    mass of one.  record=True keeps every evaluated (step, likelihood) in
    .history."""
    ctx = st.ctx
    nstep = int(ctx.nStep)
    best = BestLikelihood()
    use_disp = use_beta_disp or use_vel_disp or use_vlos or use_beta_comp or use_dist \
        or use_prop_mot
    st.gravity()
    step = st.tree_info().step
    while step < nstep:
        st.step()
        step = st.tree_info().step
        if step / float(nstep) < best_like_start:
            continue
        hist = simulated_histogram(st, hp, total_simulated, mass_per_particle,
                                   dispersions=use_disp, **disp)
        arr = L.system_likelihood(data, hist, st.n, use_beta_disp=use_beta_disp,
                                  use_vel_disp=use_vel_disp, use_vlos=use_vlos,
                                  use_beta_comp=use_beta_comp, use_dist=use_dist,
                                  use_prop_mot=use_prop_mot)
        like = clamp_likelihood(float(arr[0]))
        if record:
            best.history.append((step, like))
        if abs(like) < abs(best.likelihood):
            best.likelihood = like
            best.emd, best.mass = float(arr[1]), float(arr[2])
            best.beta = float(arr[3]) if use_beta_disp else 0.0
            best.vel = float(arr[4]) if use_vel_disp else 0.0
            best.beta_avg = float(arr[5]) if use_beta_comp else 0.0
            best.vel_avg = float(arr[6]) if use_vlos else 0.0
            best.dist = float(arr[7]) if use_dist else 0.0
            best.pm_dec = float(arr[8]) if use_prop_mot else 0.0
            best.pm_ra = float(arr[9]) if use_prop_mot else 0.0
            best.time = (float(step) / float(nstep)) * ctx.timeEvolve
            best.step = step
            best.count += 1
            best.histogram = hist
            if keep_bodies:
                best.bodies = st.bodies()
    return best
