"""Host-side lambda-beta histogram of a final state and the EMD between two
such histograms, for reporting what the step loop produces (the client's
own likelihood code stays as it is).

nbCreateHistogram's normalised counts (nbody/src/nbody_histogram.c:671-940):
bodies that are not ignored (the dark matter of a mixed dwarf is) are binned
by lambda, beta of nbXYZToLambdaBeta (nbody/src/nbody_coordinates.c:129-164)
with index floor((value - start) / size); counts are divided by the number
of bodies inside the histogram.  The EMD is that of nbMatchEMD
(nbody/src/nbody_emd.c:1278-1428) over the full lambda range for a
single-beta-bin histogram: the transport distance between the two
normalised histograms with bins at their lambda centres (degrees), rounded
to 1e-9 as the reference does.  For one beta bin this is the 1-D
Wasserstein distance, sum over bins of |CDF_a - CDF_b| times the bin width.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class HistogramParams:
    """HistogramParams.create of the sample workunits
    (nbody/sample_workunits/for_developers.lua:370-385)."""
    phi: float = 128.79
    theta: float = 54.39
    psi: float = 90.70
    lambdaStart: float = -150.0
    lambdaEnd: float = 150.0
    lambdaBins: int = 50
    betaStart: float = -15.0
    betaEnd: float = 15.0
    betaBins: int = 1
    sunGCDist: float = 8.0

    def rotation(self) -> np.ndarray:
        f, t, s = (math.radians(a) for a in (self.phi, self.theta, self.psi))
        cf, sf, ct, st, cs, ss = (math.cos(f), math.sin(f), math.cos(t), math.sin(t),
                                  math.cos(s), math.sin(s))
        return np.array([[cs * cf - ct * sf * ss, cs * sf + ct * cf * ss, ss * st],
                         [-ss * cf - ct * sf * cs, -ss * sf + ct * cf * cs, cs * st],
                         [st * sf, -st * cf, ct]])


def lambda_beta(pos: np.ndarray, hp: HistogramParams):
    """(lambda, beta) in degrees of Galactocentric positions [n, 3]."""
    p = np.array(pos, dtype=np.float64).reshape(-1, 3)
    p[:, 0] += hp.sunGCDist
    t = p @ hp.rotation().T
    r = np.sqrt((t * t).sum(1))
    lam = np.degrees(np.arctan2(t[:, 1], t[:, 0]))
    beta = np.degrees(np.arcsin(-t[:, 2] / r))
    return lam, beta


def create_histogram(pos: np.ndarray, hp: HistogramParams, ignore=None):
    """Normalised counts [lambdaBins * betaBins] (lambda-major, as the
    reference's Histindex) and the number of bodies inside."""
    if ignore is not None:
        pos = np.asarray(pos)[~np.asarray(ignore, dtype=bool)]
    lam, beta = lambda_beta(pos, hp)
    ls = (hp.lambdaEnd - hp.lambdaStart) / hp.lambdaBins
    bs = (hp.betaEnd - hp.betaStart) / hp.betaBins
    li = np.floor((lam - hp.lambdaStart) / ls)
    bi = np.floor((beta - hp.betaStart) / bs)
    ok = (li >= 0) & (li < hp.lambdaBins) & (bi >= 0) & (bi < hp.betaBins)
    idx = (li[ok] * hp.betaBins + bi[ok]).astype(np.int64)
    counts = np.bincount(idx, minlength=hp.lambdaBins * hp.betaBins).astype(np.float64)
    total = int(ok.sum())
    return (counts / total if total else counts), total


def emd(a: np.ndarray, b: np.ndarray, hp: HistogramParams) -> float:
    """EMD (degrees) between two histograms [lambdaBins * betaBins]
    (lambda-major), each normalised to unit mass, rounded to 1e-9: the
    cumulative-difference closed form for one beta bin, the transport linear
    program over the bin centres otherwise."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.sum() <= 0 or b.sum() <= 0:
        return math.inf
    if hp.betaBins != 1:
        from . import likelihood as L
        ls = (hp.lambdaEnd - hp.lambdaStart) / hp.lambdaBins
        bs = (hp.betaEnd - hp.betaStart) / hp.betaBins
        i, j = np.meshgrid(np.arange(hp.lambdaBins), np.arange(hp.betaBins), indexing="ij")
        lam = ((i + 0.5) * ls + hp.lambdaStart).ravel()
        beta = ((j + 0.5) * bs + hp.betaStart).ravel()
        d = L.emd_calc(a / a.sum(), lam, beta, b / b.sum(), lam, beta)
        return round(d * 1e9) * 1e-9
    ca = np.cumsum(a / a.sum())
    cb = np.cumsum(b / b.sum())
    w = (hp.lambdaEnd - hp.lambdaStart) / hp.lambdaBins
    d = float(np.abs(ca - cb)[:-1].sum() * w)
    return round(d * 1e9) * 1e-9
