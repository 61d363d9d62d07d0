"""Numpy restatement of the group walk's interaction lists (the checker of
precision MIXED_GROUP; TEST INFRASTRUCTURE ONLY, used by tests/ and never
by the product).

Given a tree in the reference's pre-order (kind, Next, centre of mass,
mass, rcrit2, quadrupole -- the oracle's tree, identical to the GPU's), the
bodies in Morton order and the root size, it forms the GPU walk's groups
(64 consecutive bodies; cut where the simulation changes and, twice, at the
weakest common-prefix link when its level lies 3 or more levels below the
lower median of the run's link levels), accepts for each group every node
whose centre of mass is at squared distance >= rcrit2 (times the open
scale, rounded up to FP32) from the group's bounding box, or that is a
body, opens the rest, and evaluates each body's acceleration over its
group's list in FP64 with the reference's softened monopole + quadrupole
expressions (nbody/src/nbody_grav.c:53-99).

The set of accepted nodes depends only on the tests (a node is tested iff
all its ancestors were opened), not on the order in which the GPU visits
them, so the interaction counts must agree exactly and the accelerations
to the FP32 rounding of the GPU's interaction loop.
"""
from __future__ import annotations

import numpy as np

SPLIT_T = 2
SPLIT_ABS = 4
SPLIT_PASSES = 3
WAVE = 64


def children(kind, skip):
    """Children (pre-order indices) of every node: kids[p, :8], -1 padded."""
    nn = kind.size
    kids = -np.ones((nn, 8), np.int64)
    nk = np.zeros(nn, np.int64)
    stack = []
    for i in range(nn):
        while stack and skip[stack[-1]] <= i:
            stack.pop()
        if stack:
            p = stack[-1]
            kids[p, nk[p]] = i
            nk[p] += 1
        if kind[i] == 0:
            stack.append(i)
    return kids


def _bit_length(x):
    x = x.astype(np.uint64)
    bl = np.zeros(x.shape, np.int64)
    for k in range(64):
        bl += (x >> np.uint64(k)) > 0
    return bl


def fixed_point(x, rsize):
    """nbh_fixed: floor(x * 2^63 / R) + 2^62, clamped to [0, 2^63 - 1]."""
    t = np.floor(np.asarray(x, np.float64) * (2.0 ** 63 / rsize))
    out = np.empty(t.shape, np.uint64)
    hi = t >= 2.0 ** 62
    lo = t < -(2.0 ** 62)
    mid = ~hi & ~lo
    out[hi] = np.uint64(2 ** 63 - 1)
    out[lo] = np.uint64(0)
    out[mid] = (t[mid].astype(np.int64) + (1 << 62)).astype(np.uint64)
    return out


def link_levels(pos_sorted, rsize):
    """Level of the common Morton prefix of each consecutive pair of sorted
    bodies (189-bit keys; equal keys -> 63)."""
    u = fixed_point(pos_sorted, rsize)
    x = u[:-1] ^ u[1:]
    e = [63 - _bit_length(x[:, k]) for k in range(3)]     # leading zeros in 63 bits
    m = np.minimum(e[0], np.minimum(e[1], e[2]))
    bits = np.where(m >= 63, 189, 3 * m + np.where(e[0] == m, 0, np.where(e[1] == m, 1, 2)))
    return np.minimum(bits, 189) // 3


def _split_run(lvl, a, e, cuts):
    """One pass over run [a, e] (inclusive lanes): cut after the first
    weakest link when its level is more than SPLIT_T below the lower median,
    or at most SPLIT_ABS."""
    nl = e - a
    if nl < 1:
        return
    d = lvl[a:e]
    if (d < 0).any():
        return
    j = int(np.argmin(d))
    need = (nl + 1) // 2
    med = int(np.sort(d)[need - 1])
    if d[j] < med - SPLIT_T or d[j] <= SPLIT_ABS:
        cuts.add(a + j)


def wave_groups(lvl_wave, nvalid):
    """Sub-groups [a, b) (lane ranges) of one wave; lvl_wave[i] = level of
    the link between lanes i and i+1 (-1 = hard cut)."""
    cuts = {nvalid - 1}
    cuts.update(i for i in range(nvalid - 1) if lvl_wave[i] < 0)
    for _ in range(SPLIT_PASSES):
        new = set(cuts)
        a = 0
        for e in sorted(cuts):
            _split_run(lvl_wave, a, e, new)
            a = e + 1
        cuts = new
    out, a = [], 0
    for e in sorted(cuts):
        out.append((a, e + 1))
        a = e + 1
    return out


def round_up_f32(v):
    f = np.asarray(v, np.float64).astype(np.float32)
    low = f.astype(np.float64) < v
    f[low] = np.nextafter(f[low], np.float32(np.inf))
    return f.astype(np.float64)


class GroupWalk:
    """The group walk of one simulation over a given tree."""

    def __init__(self, tree, rsize, eps2, open_scale=1.0, use_quad=True):
        self.kind = tree["kind"]
        self.skip = tree["skip"]
        self.cpos = tree["pos"]
        self.mass = tree["mass"]
        self.quad = tree["quad"] if use_quad else np.zeros_like(tree["quad"])
        self.rc2 = np.where(self.kind == 0, round_up_f32(tree["rcrit2"] * open_scale), -1.0)
        self.kids = children(self.kind, self.skip)
        self.eps2 = eps2
        leaves = np.nonzero(self.kind == 1)[0]
        self.leaf = leaves                         # pre-order index of sorted body s
        self.order = tree["ref"][leaves]           # original index of sorted body s
        self.spos = self.cpos[leaves]
        self.lvl = link_levels(self.spos, rsize)   # link s -> s + 1

    def groups(self):
        """All sub-groups as (first sorted body, end)."""
        n = self.spos.shape[0]
        out = []
        for w0 in range(0, n, WAVE):
            nv = min(WAVE, n - w0)
            lw = np.full(WAVE, -1, np.int64)
            k = min(nv - 1, self.lvl.size - w0)
            if k > 0:
                lw[:k] = self.lvl[w0:w0 + k]
            out.extend((w0 + a, w0 + b) for a, b in wave_groups(lw, nv))
        return out

    def accepted(self, a, b):
        """Accepted node list of the sub-group of sorted bodies [a, b)."""
        p = self.spos[a:b]
        lo, hi = p.min(0), p.max(0)
        front = np.array([0])
        lists = []
        while front.size:
            c = self.cpos[front]
            dd = np.maximum(np.maximum(lo - c, c - hi), 0.0)
            d2 = dd[:, 2] * dd[:, 2] + (dd[:, 1] * dd[:, 1] + dd[:, 0] * dd[:, 0])
            acc = (self.kind[front] == 1) | (d2 >= self.rc2[front])
            lists.append(front[acc])
            ch = self.kids[front[~acc]].ravel()
            front = ch[ch >= 0]
        return np.concatenate(lists)

    def forces(self, a, b, lst):
        """FP64 self-gravity of sorted bodies [a, b) from the list."""
        out = np.zeros((b - a, 3))
        q = self.quad[lst]
        cp = self.cpos[lst]
        m = self.mass[lst]
        cell = self.kind[lst] == 0
        for j in range(b - a):
            d = cp - self.spos[a + j]
            r2 = (d * d).sum(1) + self.eps2
            r = np.sqrt(r2)
            mor3 = m / r / r2
            qx = q[:, 0] * d[:, 0] + q[:, 1] * d[:, 1] + q[:, 2] * d[:, 2]
            qy = q[:, 1] * d[:, 0] + q[:, 3] * d[:, 1] + q[:, 4] * d[:, 2]
            qz = q[:, 2] * d[:, 0] + q[:, 4] * d[:, 1] + q[:, 5] * d[:, 2]
            dqd = qx * d[:, 0] + qy * d[:, 1] + qz * d[:, 2]
            r5 = np.where(cell, 1.0 / (r2 * r2 * r), 0.0)
            phi = 2.5 * r5 * dqd / r2
            qv = np.stack([qx, qy, qz], 1)
            f = (mor3 + phi)[:, None] * d - r5[:, None] * qv
            out[j] = f.sum(0)
        return out

    def counts(self, groups=None):
        """(body-body interactions without self, body-cell interactions) over
        the given sub-groups (all by default)."""
        nbb = nbc = 0
        for a, b in (groups if groups is not None else self.groups()):
            lst = self.accepted(a, b)
            ncell = int((self.kind[lst] == 0).sum())
            nbody = lst.size - ncell
            own = int(np.isin(lst, self.leaf[a:b]).sum())
            nbb += (b - a) * nbody - own
            nbc += (b - a) * ncell
        return nbb, nbc


# ---------------------------------------------------------------------------
# Far-field expansion (nbhip_set_group_expansion)
# ---------------------------------------------------------------------------
# Phi(d) = -M / s - (d^T Q d) / (2 s^5), s^2 = |d|^2 + eps2, is the softened
# potential whose gradient the reference's interaction evaluates
# (a = M d / s^3 + 2.5 (d^T Q d) d / s^7 - Q d / s^5, nbody_grav.c:53-99).
# A far cell seen from a group's centre at D contributes, at body offset u,
# a(D - u) ~ g - J u + K[u, u] / 2 with g, J and K the first three
# derivatives of Phi at D (K from the monopole only).

def _derivs(D, m, Q, eps2):
    """g (n,3), J (n,3,3), K (n,3,3,3) of each cell at D (FP64)."""
    I = np.eye(3)
    s2 = (D * D).sum(1) + eps2
    s = np.sqrt(s2)
    r3, r5 = 1.0 / (s2 * s), 1.0 / (s2 * s2 * s)
    r7, r9 = r5 / s2, r5 / (s2 * s2)
    Qm = np.empty((D.shape[0], 3, 3))
    Qm[:, 0, 0], Qm[:, 0, 1], Qm[:, 0, 2] = Q[:, 0], Q[:, 1], Q[:, 2]
    Qm[:, 1, 0], Qm[:, 1, 1], Qm[:, 1, 2] = Q[:, 1], Q[:, 3], Q[:, 4]
    Qm[:, 2, 0], Qm[:, 2, 1], Qm[:, 2, 2] = Q[:, 2], Q[:, 4], Q[:, 5]
    q = np.einsum("nij,nj->ni", Qm, D)
    dqd = (q * D).sum(1)
    DD = np.einsum("ni,nj->nij", D, D)
    g = (m * r3)[:, None] * D + (2.5 * dqd * r7)[:, None] * D - r5[:, None] * q
    J = ((m * r3)[:, None, None] * I - (3 * m * r5)[:, None, None] * DD
         - r5[:, None, None] * Qm
         + (5 * r7)[:, None, None] * (np.einsum("ni,nj->nij", q, D) + np.einsum("ni,nj->nij", D, q))
         + (2.5 * dqd * r7)[:, None, None] * I - (17.5 * dqd * r9)[:, None, None] * DD)
    sym = (np.einsum("ij,nk->nijk", I, D) + np.einsum("ik,nj->nijk", I, D)
           + np.einsum("jk,ni->nijk", I, D))
    K = (-3 * m * r5)[:, None, None, None] * sym + \
        (15 * m * r7)[:, None, None, None] * np.einsum("ni,nj,nk->nijk", D, D, D)
    return g, J, K


def taylor_field(D, m, Q, eps2, u):
    """sum over cells of g - J u + K[u,u]/2 at offsets u (k,3)."""
    g, J, K = _derivs(np.atleast_2d(D), np.atleast_1d(m), np.atleast_2d(Q), eps2)
    G, JJ, KK = g.sum(0), J.sum(0), K.sum(0)
    return G[None, :] - u @ JJ.T + 0.5 * np.einsum("ijk,nj,nk->ni", KK, u, u)


def exact_field(D, m, Q, eps2, u):
    """sum over cells of the reference interaction at d = D - u (k,3)."""
    out = np.zeros((u.shape[0], 3))
    for c in range(np.atleast_2d(D).shape[0]):
        d = np.atleast_2d(D)[c] - u
        qq = np.atleast_2d(Q)[c]
        s2 = (d * d).sum(1) + eps2
        s = np.sqrt(s2)
        qx = qq[0] * d[:, 0] + qq[1] * d[:, 1] + qq[2] * d[:, 2]
        qy = qq[1] * d[:, 0] + qq[3] * d[:, 1] + qq[4] * d[:, 2]
        qz = qq[2] * d[:, 0] + qq[4] * d[:, 1] + qq[5] * d[:, 2]
        dqd = qx * d[:, 0] + qy * d[:, 1] + qz * d[:, 2]
        r5 = 1.0 / (s2 * s2 * s)
        out += (np.atleast_1d(m)[c] / (s2 * s) + 2.5 * r5 * dqd / s2)[:, None] * d \
            - r5[:, None] * np.stack([qx, qy, qz], 1)
    return out


def _split(self, a, b, ratio):
    """(near list, far cells) of sub-group [a, b): a cell is far when the
    group's half box diagonal is at most ratio times its distance from the
    group's centre (the GPU's FP64 test, same association)."""
    lst = self.accepted(a, b)
    p = self.spos[a:b]
    lo, hi = p.min(0), p.max(0)
    g = 0.5 * (lo + hi)
    h = 0.5 * (hi - lo)
    umax2 = h[2] * h[2] + (h[1] * h[1] + h[0] * h[0])
    D = self.cpos[lst] - g
    dist2 = D[:, 2] * D[:, 2] + (D[:, 1] * D[:, 1] + D[:, 0] * D[:, 0])
    far = (self.kind[lst] == 0) & (umax2 <= (ratio * ratio) * dist2) & (ratio > 0)
    return lst[~far], lst[far]


def _expanded_forces(self, a, b, ratio):
    """FP64 self-gravity of sorted bodies [a, b) with far cells expanded."""
    near, far = self.split(a, b, ratio)
    out = self.forces(a, b, near)
    if far.size:
        p = self.spos[a:b]
        g = 0.5 * (p.min(0) + p.max(0))
        out = out + taylor_field(self.cpos[far] - g, self.mass[far], self.quad[far], self.eps2,
                                 p - g)
    return out


def _expanded_counts(self, ratio, groups=None):
    """(body-body without self, near body-cell, far (cell, group) pairs)."""
    nbb = nbc = nfar = 0
    for a, b in (groups if groups is not None else self.groups()):
        near, far = self.split(a, b, ratio)
        ncell = int((self.kind[near] == 0).sum())
        nbody = near.size - ncell
        own = int(np.isin(near, self.leaf[a:b]).sum())
        nbb += (b - a) * nbody - own
        nbc += (b - a) * ncell
        nfar += far.size
    return nbb, nbc, nfar


GroupWalk.split = _split
GroupWalk.expanded_forces = _expanded_forces
GroupWalk.expanded_counts = _expanded_counts
