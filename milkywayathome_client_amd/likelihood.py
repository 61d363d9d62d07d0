"""Likelihood of a simulated lambda-beta histogram against a data histogram:
the output layer that consumes the GPU histogram (nbhip_histogram).

* `read_histogram` / `write_histogram`: the reference's .hist text format
  (nbReadHistogram, nbody/src/nbody_histogram.c:1140-1570; nbPrintHistogram,
  :515-600).
* `from_counts`: a simulated histogram from raw bin counts, normalised as
  nbNormalizeHistogram does (:621-660) with the binomial count error of
  nbNormalizedHistogramError (:60-64).
* `emd_calc`: the earth mover's distance between two weighted point sets
  with the Euclidean (lambda, beta) ground distance (emdCalc,
  nbody/src/nbody_emd.c:1201-1268): exact 1-D transport when every point
  shares one beta, a linear program otherwise.
* `match_emd` (nbMatchEMD, nbody/src/nbody_emd.c:1277-1428), `cost_component`
  (nbCostComponent, nbody/src/nbody_mass.c:318-392), `nb_likelihood`
  (nbLikelihood, :395-470, the comparison of a per-bin quantity such as a
  dispersion, averaged over betaDispBins bins) and `system_likelihood`
  (nbSystemLikelihood, nbody/src/nbody_likelihood.c:155-323): the geometry,
  mass and dispersion terms, with the reference's NaN / infinity conventions.

The per-bin histograms of a simulation (beta and line-of-sight velocity
dispersions, average line-of-sight velocity, beta, distance and proper
motions) come from the GPU (nbhip_histogram_dispersion, with the reference's
outlier rejection).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .histogram import HistogramParams

DEFAULT_WORST_CASE = 9999999.9          # nbody/include/nbody_defaults.h:77


def c_round(x):
    """C round(): halves away from zero (numpy rounds halves to even)."""
    x = np.asarray(x, dtype=np.float64)
    a = np.abs(x)
    r = np.floor(a)
    return np.sign(x) * (r + (a - r >= 0.5))


@dataclass
class Histogram:
    """The counts histogram of a MainStruct (histograms[0]) plus the header
    fields the likelihood reads."""
    lambda_bins: int
    beta_bins: int
    total_num: int                      # "n": bodies inside the binned range
    total_simulated: int                # bodies that could be binned
    mass_per_particle: float
    use_bin: np.ndarray                 # [nbin] int
    lam: np.ndarray                     # [nbin] bin centre
    beta: np.ndarray                    # [nbin] bin centre
    variable: np.ndarray                # [nbin] normalised counts
    err: np.ndarray                     # [nbin] count error
    emd_range: List[float] = field(default_factory=list)
    beta_disp_bins: int = 0
    extra: Optional[np.ndarray] = None  # [nbin, k] remaining columns, as read
    # histograms 1.. of the MainStruct: k -> (variable, err) per bin
    # (1 beta dispersion, 2 line-of-sight velocity dispersion, 3 LOS velocity,
    # 4 beta average, 5 distance, 6 / 7 proper motions)
    comps: dict = field(default_factory=dict)

    @property
    def nbin(self) -> int:
        return self.lambda_bins * self.beta_bins


def _emd_range_tokens(s: str) -> List[float]:
    for ch in "{}()[],":
        s = s.replace(ch, " ")
    # strtof: each bound is rounded to single precision
    return [float(np.float32(t)) for t in s.split()]


def read_histogram(path: str) -> Histogram:
    """Parse a .hist file as nbReadHistogram does: comment and blank lines
    skipped, each header key read once, a '!' line announcing the 15-column
    (19 with hasPM) row format, otherwise 9 columns."""
    hdr = {}
    rows = []
    orbit_cols = False
    with open(path) as f:
        lines = f.readlines()
    for raw in lines:
        line = raw.rstrip("\n")
        if not line or line.startswith("#"):
            continue
        if line.startswith("!"):
            orbit_cols = True
            continue
        t = line.strip()
        if t in ("<histogram>", "</histogram>"):
            continue
        if "=" in t:
            key, _, val = t.partition("=")
            key = key.strip()
            if key not in hdr:
                hdr[key] = val.strip()
            continue
        rows.append(t.split())
    lb = int(hdr["lambdaBins"])
    bb = int(hdr["betaBins"])
    ncol = (19 if int(hdr.get("hasPM", "0")) == 1 else 15) if orbit_cols else 9
    # like the reference's sscanf, columns beyond the format's are ignored
    if len(rows) != lb * bb or any(len(r) < ncol for r in rows):
        raise ValueError(f"{path}: expected {lb * bb} rows of {ncol} columns")
    a = np.array([[float(x) for x in r[:ncol]] for r in rows])
    h = Histogram(lambda_bins=lb, beta_bins=bb, total_num=int(hdr.get("n", "0")),
                  total_simulated=int(hdr.get("totalSimulated", "0")),
                  mass_per_particle=float(hdr.get("massPerParticle", "0")),
                  use_bin=a[:, 0].astype(np.int64), lam=a[:, 1].copy(), beta=a[:, 2].copy(),
                  variable=a[:, 3].copy(), err=a[:, 4].copy(),
                  beta_disp_bins=int(hdr.get("betaDispBins", "0")), extra=a[:, 5:].copy())
    for k in range(1, (ncol - 5) // 2 + 1):
        h.comps[k] = (a[:, 3 + 2 * k].copy(), a[:, 4 + 2 * k].copy())
    if "EMDRange" in hdr:
        h.emd_range = _emd_range_tokens(hdr["EMDRange"])
    else:
        # the whole histogram, half a bin beyond the first and the
        # lambdaBins-th centre (nbody_histogram.c, end of nbReadHistogram)
        w = (h.lam[1] - h.lam[0]) / 2.0 if h.nbin > 1 else 0.0
        h.emd_range = [h.lam[0] - w, h.lam[lb - 1] + w]
    return h


def write_histogram(path: str, h: Histogram, comment: str = "") -> None:
    """nbPrintHistogram (nbody/src/nbody_histogram.c:515-600): the counts and
    their errors, then (value, error) of histograms 1-2 (9 columns), or of
    1-7 behind a '!' line when histogram 3 is present (19 columns; 'hasPM = 1'
    when histogram 6 is); -1 where a histogram was not computed."""
    with open(path, "w") as f:
        for c in comment.splitlines():
            f.write(f"# {c}\n")
        f.write("<histogram>\n")
        f.write(f"n = {h.total_num}\n")
        f.write(f"massPerParticle = {h.mass_per_particle:12.15f}\n")
        f.write(f"totalSimulated = {h.total_simulated}\n")
        f.write(f"lambdaBins = {h.lambda_bins}\n")
        f.write(f"betaBins = {h.beta_bins}\n")
        if h.beta_disp_bins:
            f.write(f"betaDispBins = {h.beta_disp_bins}\n")
        orbit = 3 in h.comps
        if 6 in h.comps:
            f.write("hasPM = 1\n")
        if orbit:
            f.write("!\n")
        for i in range(h.nbin):
            cols = []
            for k in (range(1, 8) if orbit else (1, 2)):
                if k in h.comps:
                    cols += [f"{h.comps[k][0][i]:12.15f}", f"{h.comps[k][1][i]:12.15f}"]
                else:
                    cols += ["-1", "-1"]
            f.write(f"{int(h.use_bin[i])} {h.lam[i]:12.15f} {h.beta[i]:12.15f} "
                    f"{h.variable[i]:12.15f} {h.err[i]:12.15f} " + " ".join(cols) + "\n")
        f.write("</histogram>\n")


def normalized_error(n: int, total_histogram: float, total_simulated: float) -> float:
    """Binomial error of a normalised count (nbNormalizedHistogramError)."""
    if n == 0:
        return math.sqrt((1.0 / total_simulated) * (1.0 - 1.0 / total_simulated)) / total_histogram
    rn = float(n)
    return math.sqrt(rn * (rn / total_simulated) * (1.0 - rn / total_simulated)) / total_histogram


def from_counts(counts, hp: HistogramParams, total_simulated: int,
                mass_per_particle: float, dispersions=None) -> Histogram:
    """A simulated histogram from raw counts [lambdaBins * betaBins]
    (lambda-major, as nbhip_histogram returns them): bin centres, normalised
    counts and errors as nbNormalizeHistogram computes them.  Like the
    reference, the error of bin (i, j) is computed from the raw count of bin
    i (identical for one beta bin).  `dispersions` = {k: (value, error)} for
    histograms 1-7, as NBodyState.histogram_dispersion returns it."""
    c = np.asarray(counts, dtype=np.int64).reshape(-1)
    lb, bb = hp.lambdaBins, hp.betaBins
    if c.size != lb * bb:
        raise ValueError("counts do not match the histogram's bins")
    total = int(c.sum())
    ls = (hp.lambdaEnd - hp.lambdaStart) / lb
    bs = (hp.betaEnd - hp.betaStart) / bb
    lam = np.empty(c.size)
    beta = np.empty(c.size)
    var = np.empty(c.size)
    err = np.empty(c.size)
    for i in range(lb):
        for j in range(bb):
            k = i * bb + j
            lam[k] = (i + 0.5) * ls + hp.lambdaStart
            beta[k] = (j + 0.5) * bs + hp.betaStart
            var[k] = c[k] / total if total else math.nan
            err[k] = normalized_error(int(c[i]), total, total_simulated) if total else math.nan
    h = Histogram(lambda_bins=lb, beta_bins=bb, total_num=total,
                  total_simulated=int(total_simulated),
                  mass_per_particle=float(mass_per_particle),
                  use_bin=np.ones(c.size, dtype=np.int64), lam=lam, beta=beta, variable=var,
                  err=err, beta_disp_bins=1)
    if dispersions is not None:
        for k, (v, e) in dispersions.items():
            h.comps[int(k)] = (np.asarray(v, dtype=np.float64).reshape(-1),
                               np.asarray(e, dtype=np.float64).reshape(-1))
    return h


# Largest transport problem emd_calc solves as a linear program (sparse
# constraints: memory grows with the number of flows, (lambda x beta bins)^2).
EMD_LP_MAX_FLOWS = 4_000_000


def emd_calc(w1, lam1, beta1, w2, lam2, beta2) -> float:
    """Minimum total transport cost over the total flow, min(sum w1, sum w2),
    with ground distance sqrt(dlambda^2 + dbeta^2) (emdCalc).  NaN when a
    weight is not finite or a side carries no weight."""
    w1 = np.asarray(w1, dtype=np.float64)
    w2 = np.asarray(w2, dtype=np.float64)
    lam1, lam2 = np.asarray(lam1, np.float64), np.asarray(lam2, np.float64)
    beta1, beta2 = np.asarray(beta1, np.float64), np.asarray(beta2, np.float64)
    if w1.size == 0 or w2.size == 0 or not (np.all(np.isfinite(w1)) and np.all(np.isfinite(w2))):
        return math.nan
    s1, s2 = w1.sum(), w2.sum()
    flow = min(s1, s2)
    if not flow > 0.0:
        return math.nan
    one_beta = np.all(beta1 == beta1[0]) and np.all(beta2 == beta1[0])
    if one_beta and abs(s1 - s2) <= 1e-12 * max(s1, s2):
        # 1-D transport of equal masses: integral of |F1 - F2| over lambda
        x = np.concatenate([lam1, lam2])
        order = np.argsort(x, kind="stable")
        xs = x[order]
        d = np.concatenate([w1 / s1, -w2 / s2])[order]
        cdf = np.cumsum(d)[:-1]
        return float(np.sum(np.abs(cdf) * np.diff(xs)))
    from scipy import sparse
    from scipy.optimize import linprog
    n1, n2 = w1.size, w2.size
    if n1 * n2 > EMD_LP_MAX_FLOWS:
        raise ValueError(f"emd_calc: {n1} x {n2} bins is {n1 * n2} transport variables, "
                         f"above the {EMD_LP_MAX_FLOWS} this solver takes")
    cost = np.sqrt((lam1[:, None] - lam2[None, :]) ** 2 + (beta1[:, None] - beta2[None, :]) ** 2)
    # flow f[i, j] at column i * n2 + j: row i sums the flows out of source bin
    # i, row n1 + j the flows into target bin j (two nonzeros per column)
    cols = np.arange(n1 * n2)
    rows = np.concatenate([cols // n2, n1 + cols % n2])
    a_ub = sparse.csr_matrix((np.ones(2 * cols.size), (rows, np.concatenate([cols, cols]))),
                             shape=(n1 + n2, n1 * n2))
    a_eq = sparse.csr_matrix(np.ones((1, n1 * n2)))
    res = linprog(cost.ravel(), A_ub=a_ub, b_ub=np.concatenate([w1, w2]),
                  A_eq=a_eq, b_eq=[flow], bounds=(0, None), method="highs")
    if not res.success:
        return math.nan
    return float(res.fun / flow)


def match_emd(data: Histogram, hist: Histogram) -> float:
    """nbMatchEMD: -300 log(1 - EMD / 50), the EMD taken over each of the
    data histogram's lambda ranges (renormalised inside the range, bins with
    useBin = 0 left out) and weighted by the data counts in the range; NaN
    for mismatched bins, missing masses or an EMD above 50, infinity for an
    empty histogram."""
    if data.lambda_bins != hist.lambda_bins or data.beta_bins != hist.beta_bins:
        return math.nan
    n_sim = int(hist.total_num)
    n_data = int(data.total_num)
    if n_sim == 0 or n_data == 0:
        return math.inf
    if hist.mass_per_particle <= 0.0 or data.mass_per_particle <= 0.0:
        return math.nan
    rng = list(data.emd_range)
    if len(rng) % 2:
        return math.nan
    if not rng:
        rng = [data.lam[0], data.lam[data.nbin - 1]]
    emd = 0.0
    total_range = 0
    use = data.use_bin != 0
    for r in range(0, len(rng), 2):
        start, end = rng[r], rng[r + 1]
        if start > end:
            return math.nan
        sel_h = (hist.lam >= start) & (hist.lam <= end) & use
        range_h = int(np.sum(c_round(hist.variable[sel_h] * n_sim)))
        wh = hist.variable[sel_h] * n_sim / (1.0 * range_h) if range_h else \
            np.full(int(sel_h.sum()), math.nan)
        sel_d = (data.lam >= start) & (data.lam <= end) & use
        range_d = int(np.sum(c_round(data.variable[sel_d] * n_data)))
        total_range += range_d
        wd = data.variable[sel_d] * n_data / (1.0 * range_d) if range_d else \
            np.full(int(sel_d.sum()), math.nan)
        e = emd_calc(wd, data.lam[sel_d], data.beta[sel_d], wh, hist.lam[sel_h], hist.beta[sel_h])
        emd += e * (range_d / (1.0 * n_data))
    if total_range == 0:
        return math.nan
    # the reference scales by an unsigned integer quotient
    emd = emd * (n_data // total_range)
    emd = float(c_round(emd * 1.0e9)) * 1.0e-9
    if not emd <= 50.0:
        return math.nan
    return -300.0 * math.log(1.0 - emd / 50.0)


def cost_component(data: Histogram, hist: Histogram) -> float:
    """nbCostComponent: minus the log of the Gaussian probability that the
    simulated baryon mass in the used bins matches the data's, with the
    data's count errors and the simulation's binomial spread."""
    if data.lambda_bins != hist.lambda_bins or data.beta_bins != hist.beta_bins:
        return math.nan
    n = float(hist.total_simulated)
    n_sim = float(hist.total_num)
    n_data = float(data.total_num)
    if n_sim == 0.0 or n_data == 0.0:
        return math.inf
    hm, dm = hist.mass_per_particle, data.mass_per_particle
    if hm <= 0.0 or dm <= 0.0:
        return math.nan
    var_data = 0.0
    for i in range(data.nbin):
        if not data.use_bin[i]:
            n_sim -= float(c_round(hist.variable[i] * float(hist.total_num)))
        else:
            var_data += (data.err[i] * n_data) ** 2
    p = n_sim / n
    num = -(dm * n_data - hm * n_sim) ** 2
    denom = 2.0 * (dm * dm * var_data + hm * hm * n_sim * p * (1.0 - p))
    return -(num / denom)


def nb_likelihood(data_var, data_err, use_bin, hist_var, hist_err, avg_bins: int) -> float:
    """nbLikelihood: half the sum over used data bins of
    (data - sim)^2 / (err_data^2 + err_sim^2); 25 (five sigma) where the
    simulation has no error; bins whose data error is not positive are
    skipped.  With avg_bins > 1 the simulated value is the average over the
    avg_bins bins centred on the data bin (an even count is reduced by one),
    leaving out bins outside the histogram or without an error."""
    data_var = np.asarray(data_var, dtype=np.float64)
    data_err = np.asarray(data_err, dtype=np.float64)
    hist_var = np.asarray(hist_var, dtype=np.float64)
    hist_err = np.asarray(hist_err, dtype=np.float64)
    nbins = data_var.size
    nsig = 0.0
    for i in range(nbins):
        if not use_bin[i]:
            continue
        if avg_bins > 1:
            if avg_bins % 2 == 0:
                avg_bins -= 1
            n = avg_bins
            var_sum = 0.0
            val_sum = 0.0
            for k in range(avg_bins):
                idx = i - (avg_bins - 1) // 2 + k
                if idx < 0 or idx >= nbins:
                    n -= 1
                elif hist_err[idx] <= 0:
                    n -= 1
                else:
                    var_sum += hist_err[idx] ** 2
                    val_sum += hist_var[idx]
            with np.errstate(divide="ignore", invalid="ignore"):
                err_hist = float(np.sqrt(var_sum) / n) if n else math.nan
                hist_v = val_sum / n if n else math.nan
            err_data = data_err[i]
            data_v = data_var[i]
        else:
            err_data = data_err[i]
            err_hist = hist_err[i]
            data_v = data_var[i]
            hist_v = hist_var[i]
        if err_data > 0:
            if err_hist > 0:
                nsig += (data_v - hist_v) ** 2 / (err_data ** 2 + err_hist ** 2)
            else:
                nsig += 25.0
    return nsig / 2.0


def worst_case_emd() -> float:
    return DEFAULT_WORST_CASE


def system_likelihood(data: Histogram, hist: Histogram, nbody: int, method: str = "emd",
                      use_beta_disp: bool = False, use_vel_disp: bool = False,
                      use_vlos: bool = False, use_beta_comp: bool = False,
                      use_dist: bool = False, use_prop_mot: bool = False) -> np.ndarray:
    """nbSystemLikelihood: [likelihood, geometry, cost, beta dispersion,
    velocity dispersion, beta, LOS velocity, distance, dec PM, RA PM], the
    unused terms NaN; all NaN for mismatched bins; the worst-case EMD in
    the first two entries when fewer than 1e-4 of the bodies are binned."""
    out = np.full(10, math.nan)
    if data.lambda_bins != hist.lambda_bins:
        return out
    # (flag, histogram, output slot) in the reference's order of addition
    terms = ((use_beta_disp, 1, 3), (use_vel_disp, 2, 4), (use_vlos, 3, 6),
             (use_beta_comp, 4, 5), (use_dist, 5, 7), (use_prop_mot, 6, 8),
             (use_prop_mot, 7, 9))
    for flag, k, _ in terms:
        if flag and (k not in data.comps or k not in hist.comps):
            raise ValueError(f"histogram {k} missing from the data or the simulation "
                             "(from_counts(..., dispersions=...))")
    if method != "emd":
        # nbCalcChisq (nbody/src/nbody_chisq.c:131-232) iterates lambdaBins^2
        # bins over lambdaBins * betaBins entries: no defined result to match
        raise NotImplementedError("only the EMD geometry term is implemented")
    if hist.total_num < 0.0001 * float(nbody):
        out[:] = 0.0
        out[0] = out[1] = worst_case_emd()
        return out
    geometry = match_emd(data, hist)
    cost = cost_component(data, hist)
    like = geometry + cost
    # [likelihood, geometry, cost, beta disp, vel disp, beta, LOS velocity,
    # distance, dec PM, RA PM]: the reference's order of components
    for flag, k, slot in terms:
        if flag:
            avg = data.beta_disp_bins if k == 1 else 1
            out[slot] = nb_likelihood(*data.comps[k], data.use_bin, *hist.comps[k], avg)
            like += out[slot]
    out[0] = like
    out[1] = geometry
    out[2] = cost
    return out
