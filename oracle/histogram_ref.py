"""Test infrastructure only (the checker; the product never imports this):
a numpy restatement of nbCreateHistogram's per-bin histograms 1-7 (beta and
line-of-sight velocity dispersions; average line-of-sight velocity, beta,
distance and proper motions in declination and right ascension) with the
reference's outlier rejection, summed in body order as the reference does.

  * binning and (lambda, beta): nbody/src/nbody_histogram.c:858-903, through
    the host lambda_beta (milkywayathome_client_amd/histogram.py, pinned by
    tests/test_histogram.py);
  * calc_vLOS and calc_distance: nbody/src/nbody_mass.c:171-187;
  * nbVXVYVZtomuDec / nbVXVYVZtomuRA with cartesianalign, cartesianToLbr_rad,
    findLatUnitVec and findLongUnitVec: nbody/src/nbody_coordinates.c:24-39,
    166-287 (defaults nbody/include/nbody_defaults.h:40-45);
  * per-bin sums: nbody/src/nbody_histogram.c:905-917 (np.add.at adds in
    index order, i.e. body order);
  * nbCalcDisp: nbody/src/nbody_mass.c:195-245 (only bins with more than 10
    bodies change; the correction factor on every call after the first);
  * nbRemoveOutliers: nbody/src/nbody_mass.c:247-316 (mean and dispersion of
    the previous round, keep |mean - value| < sigma * dispersion);
  * the rounds: nbody/src/nbody_histogram.c:964-999; the averages and their
    errors (dispersion / sqrt(n), -1 and 0 below 11 bodies): :996-1110.

Pinned by closed forms in tests/test_dispersion_ref.py (symmetric bins,
a planted outlier, bins of 10 or fewer bodies, the line-of-sight velocity of
a radial orbit; the frame rotation by where it sends the Galactic and
celestial poles, the proper motions by finite differences of the equatorial
angles along the motion).
"""
import numpy as np

from milkywayathome_client_amd import histogram as H


def body_values(pos, vel, hp: H.HistogramParams, ignore=None, distance=False):
    """Bin index (-1: ignored or outside), beta (degrees), v_los per body
    (and the heliocentric distance with distance=True)."""
    pos = np.asarray(pos, dtype=np.float64)
    vel = np.asarray(vel, dtype=np.float64)
    lam, beta = H.lambda_beta(pos, hp)
    ls = (hp.lambdaEnd - hp.lambdaStart) / hp.lambdaBins
    bs = (hp.betaEnd - hp.betaStart) / hp.betaBins
    li = np.floor((lam - hp.lambdaStart) / ls)
    bi = np.floor((beta - hp.betaStart) / bs)
    ok = (li >= 0) & (li < hp.lambdaBins) & (bi >= 0) & (bi < hp.betaBins)
    if ignore is not None:
        ok &= ~np.asarray(ignore, dtype=bool)
    b = np.where(ok, li * hp.betaBins + bi, -1).astype(np.int64)
    xs = pos[:, 0] + hp.sunGCDist
    mag = np.sqrt(xs * xs + pos[:, 1] ** 2 + pos[:, 2] ** 2)
    vlos = (xs * vel[:, 0] + pos[:, 1] * vel[:, 1] + pos[:, 2] * vel[:, 2]) / mag
    if distance:
        return b, beta, vlos, mag
    return b, beta, vlos


SUN_VEL = (10.3, 229.2, 6.9)
NGP_DEC, NGP_RA, L_NCP = 27.4, 192.0, 123.0       # degrees


def galactic_to_equatorial(ngp_dec=NGP_DEC, ngp_ra=NGP_RA, l_ncp=L_NCP):
    """cartesianalign's rotation (nbody/src/nbody_coordinates.c:192-224), the
    angles in degrees."""
    d2r = np.pi / 180.0
    ln, dn, an = l_ncp * d2r, ngp_dec * d2r, ngp_ra * d2r
    cl, sl = np.cos(np.pi - ln), np.sin(np.pi - ln)
    cd, sd = np.cos(np.pi / 2 - dn), np.sin(np.pi / 2 - dn)
    ca, sa = np.cos(an), np.sin(an)
    return np.array([[ca * cd * cl - sa * sl, -ca * cd * sl - sa * cl, ca * sd],
                     [sa * cd * cl + ca * sl, -sa * cd * sl + ca * cl, sa * sd],
                     [-sd * cl, sd * sl, cd]])


def proper_motions(pos, vel, sun_gc_dist, sun_vel=SUN_VEL, ngp_dec=NGP_DEC, ngp_ra=NGP_RA,
                   l_ncp=L_NCP):
    """(mu_dec, mu_ra) per body in milliarcseconds per year of the model's
    units (nbVXVYVZtomuDec / RA, nbody/src/nbody_coordinates.c:226-287)."""
    rot = galactic_to_equatorial(ngp_dec, ngp_ra, l_ncp)
    p = np.asarray(pos, dtype=np.float64) + np.array([sun_gc_dist, 0.0, 0.0])
    u = np.asarray(vel, dtype=np.float64) - np.asarray(sun_vel, dtype=np.float64)
    P = p @ rot.T
    V = u @ rot.T
    lon = np.arctan2(P[:, 1], P[:, 0])
    lat = np.arctan2(P[:, 2], np.sqrt(P[:, 0] ** 2 + P[:, 1] ** 2))
    lon = np.where(lon < 0.0, lon + 2.0 * np.pi, lon)
    b2 = (90.0 - lat * 180.0 / np.pi) * np.pi / 180.0
    e_dec = np.stack([-np.cos(lon) * np.cos(b2), -np.cos(b2) * np.sin(lon), np.sin(b2)], 1)
    e_ra = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)], 1)
    dist = np.sqrt(np.sum(P * P, axis=1))
    conv = 180.0 / np.pi * 3600.0 / 1e6
    return (np.sum(V * e_dec, axis=1) / dist * conv, np.sum(V * e_ra, axis=1) / dist * conv)


def _calc_disp(sum_, sq, removed, raw, initial, correction, var, err):
    for i in range(raw.size):
        count = raw[i] - removed[i]
        if count > 10.0:
            n_new = count - 1.0
            n_ratio = count / n_new
            dispsq = (sq[i] / n_new) - n_ratio * (sum_[i] / count) ** 2
            if not initial:
                dispsq *= correction
            # a spread far below the mean can cancel to a tiny negative: NaN,
            # as the reference's sqrt gives
            with np.errstate(invalid="ignore"):
                var[i] = np.sqrt(dispsq)
            err[i] = np.sqrt((count + 1) / (count * n_new)) * var[i]


def _sums(b, v, nbin, keep):
    s = np.zeros(nbin)
    q = np.zeros(nbin)
    r = np.zeros(nbin)
    sel = b >= 0
    np.add.at(s, b[sel & keep], v[sel & keep])
    np.add.at(q, b[sel & keep], v[sel & keep] ** 2)
    np.add.at(r, b[sel & ~keep], 1.0)
    return s, q, r


def dispersion(b, v, nbin, iter_max, sigma, correction, average=False):
    """(dispersion, error) per bin for the values v of binned bodies b; with
    average=True the average and its error instead."""
    raw = np.bincount(b[b >= 0], minlength=nbin).astype(np.float64)
    var = np.zeros(nbin)
    err = np.zeros(nbin)
    s, q, r = _sums(b, v, nbin, np.ones(b.size, dtype=bool))
    _calc_disp(s, q, r, raw, True, correction, var, err)
    for _ in range(iter_max):
        with np.errstate(divide="ignore", invalid="ignore"):
            ave = s / (raw - r)
        bb = np.where(b >= 0, b, 0)
        keep = np.abs(ave[bb] - v) < sigma * var[bb]
        s, q, r = _sums(b, v, nbin, keep)
        _calc_disp(s, q, r, raw, False, correction, var, err)
    if average:
        denom = (raw - r).astype(np.int64)
        big = denom > 10
        avg = np.where(big, s / np.where(big, denom, 1), 0.0)
        aerr = np.where(big, var / np.sqrt(np.where(big, denom, 1)), -1.0)
        return avg, aerr
    return var, err


def dispersions(pos, vel, hp, ignore=None, iter_max=6, beta_sigma=2.5, vel_sigma=2.5,
                dist_sigma=2.5, pm_sigma=2.5, beta_correct=1.111, vel_correct=1.111,
                dist_correct=1.111, pm_correct=1.111, sun_vel=SUN_VEL, ngp_dec=NGP_DEC,
                ngp_ra=NGP_RA, l_ncp=L_NCP):
    """{1: beta dispersion, 2: v_los dispersion, 3: v_los average, 4: beta
    average, 5: distance average, 6: mu_dec average, 7: mu_ra average}, each
    (value, error) per bin, as nbhip_histogram_dispersion returns them."""
    b, beta, vlos, dist = body_values(pos, vel, hp, ignore, distance=True)
    mud, mur = proper_motions(pos, vel, hp.sunGCDist, sun_vel, ngp_dec, ngp_ra, l_ncp)
    nbin = hp.lambdaBins * hp.betaBins
    return {1: dispersion(b, beta, nbin, iter_max, beta_sigma, beta_correct),
            2: dispersion(b, vlos, nbin, iter_max, vel_sigma, vel_correct),
            3: dispersion(b, vlos, nbin, iter_max, vel_sigma, vel_correct, average=True),
            4: dispersion(b, beta, nbin, iter_max, beta_sigma, beta_correct, average=True),
            5: dispersion(b, dist, nbin, iter_max, dist_sigma, dist_correct, average=True),
            6: dispersion(b, mud, nbin, iter_max, pm_sigma, pm_correct, average=True),
            7: dispersion(b, mur, nbin, iter_max, pm_sigma, pm_correct, average=True)}
