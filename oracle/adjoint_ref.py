"""CPU reference for the adjoint leapfrog: the FP64 direct-sum KDK step with
the external potential, written in numpy, and its adjoint derived by hand.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
adjoint (nbhip_adjoint_*), never by the product.

What it restates: nbGravity_Exact (nbody/src/nbody_grav.c:205-227, with the
self term, which adds 0 when eps2 > 0), nbExtAcceleration
(nbody/src/nbody_potential.c:627-762: Miyamoto-Nagai disks, logarithmic and
NFW halos, Hernquist and Plummer spheres, r clamped at 2^-8), and the KDK
step of nbStepSystemPlain (nbody/src/nbody_plain.c:268-364).  The adjoint has
no counterpart in the reference.  How it is pinned (tests/test_adjoint_ref.py):
the forward map against the oracle's Exact run; ext_jac against central
differences of ext_accel; the whole adjoint against central differences of
the forward map in every input (positions, velocities, masses, timestep).
"""
from __future__ import annotations

import numpy as np

LIMIT = 0.00390625        # r clamp of nbExtAcceleration


def pot_params(pot) -> dict:
    """Fields of an oracle OrcPotential (or None) as a plain dict."""
    if pot is None:
        return {"sphere": 0, "disk": 0, "disk2": 0, "halo": 0}
    return {"sphere": pot.sphere_type, "sm": pot.sphere_mass, "ss": pot.sphere_scale,
            "disk": pot.disk_type, "dm": pot.disk_mass, "da": pot.disk_a, "db": pot.disk_b,
            "disk2": pot.disk2_type, "d2m": pot.disk2_mass, "d2a": pot.disk2_a,
            "d2b": pot.disk2_b,
            "halo": pot.halo_type, "hv": pot.halo_vhalo, "hd": pot.halo_d, "hq": pot.halo_q}


def _radius(x):
    ab = np.sqrt((x * x).sum(1))
    return np.where(ab <= LIMIT, LIMIT, ab), ab <= LIMIT


def ext_accel(p: dict, x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
    r, _ = _radius(x)
    a = np.zeros_like(x)
    for key, m, da, db in (("disk", "dm", "da", "db"), ("disk2", "d2m", "d2a", "d2b")):
        if p[key] == 1:
            M, A, B = p[m], p[da], p[db]
            zp = np.sqrt(Z * Z + B * B)
            azp = A + zp
            rth = (X * X + Y * Y + azp * azp) ** 1.5
            a += np.stack([-M * X / rth, -M * Y / rth, -M * Z * azp / (zp * rth)], 1)
    if p["halo"] == 1:
        v0, d, q = p["hv"], p["hd"], p["hq"]
        k = -2.0 * v0 * v0 / (d * d + X * X + Y * Y + (Z / q) ** 2)
        a += np.stack([k * X, k * Y, k * Z / (q * q)], 1)
    elif p["halo"] == 2:
        h = p["hd"]
        M = p["hv"] ** 2 * h / 0.2162165954
        c = (-M / (r * r)) * (np.log((h + r) / h) / r - 1.0 / (h + r))
        a += c[:, None] * x
    if p["sphere"] == 1:
        s = -p["sm"] / (r * (p["ss"] + r) ** 2)
        a += s[:, None] * x
    elif p["sphere"] == 2:
        s = -p["sm"] / np.sqrt(p["ss"] ** 2 + r * r) ** 3
        a += s[:, None] * x
    return a


def ext_jac(p: dict, x: np.ndarray, g: np.ndarray) -> np.ndarray:
    """J g, J = d ext_accel / dx (symmetric)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    g = np.asarray(g, dtype=np.float64).reshape(-1, 3)
    X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
    r, clamp = _radius(x)
    out = np.zeros_like(x)
    for key, m, da, db in (("disk", "dm", "da", "db"), ("disk2", "d2m", "d2a", "d2b")):
        if p[key] == 1:
            M, A, B = p[m], p[da], p[db]
            zeta = np.sqrt(Z * Z + B * B)
            D = X * X + Y * Y + (A + zeta) ** 2
            u = np.stack([X, Y, Z * (A + zeta) / zeta], 1)
            ug = (u * g).sum(1)
            diag = np.stack([np.ones_like(Z), np.ones_like(Z), 1.0 + A * B * B / zeta ** 3], 1)
            out += -M * (diag * g / D[:, None] ** 1.5 - 3.0 * u * (ug / D ** 2.5)[:, None])
    if p["halo"] == 1:
        v0, d, q = p["hv"], p["hd"], p["hq"]
        D = d * d + X * X + Y * Y + (Z / q) ** 2
        k = -2.0 * v0 * v0 / D
        w = np.stack([X, Y, Z / (q * q)], 1)
        scale = np.array([1.0, 1.0, 1.0 / (q * q)])
        out += k[:, None] * scale * g - (2.0 * k / D * (w * g).sum(1))[:, None] * w
    f = np.zeros_like(r)
    fp = np.zeros_like(r)
    if p["halo"] == 2:
        h = p["hd"]
        M = p["hv"] ** 2 * h / 0.2162165954
        L = np.log((h + r) / h)
        f += -M * (L / r ** 3 - 1.0 / (r * r * (h + r)))
        fp += -M * (1.0 / (r ** 3 * (h + r)) - 3.0 * L / r ** 4
                    + (2.0 * h + 3.0 * r) / (r ** 3 * (h + r) ** 2))
    if p["sphere"] == 1:
        a = p["ss"]
        f += -p["sm"] / (r * (a + r) ** 2)
        fp += p["sm"] * (a + 3.0 * r) / (r * r * (a + r) ** 3)
    elif p["sphere"] == 2:
        s2 = p["ss"] ** 2 + r * r
        f += -p["sm"] / s2 ** 1.5
        fp += 3.0 * p["sm"] * r / s2 ** 2.5
    c = np.where(clamp, 0.0, fp / r * (x * g).sum(1))
    out += f[:, None] * g + c[:, None] * x
    return out


def direct_accel(x, m, eps2):
    d = x[None, :, :] - x[:, None, :]                 # d[i, k] = x_k - x_i
    s2 = (d * d).sum(2) + eps2
    return ((m[None, :] / np.sqrt(s2) / s2)[:, :, None] * d).sum(1)


def direct_adj(x, m, eps2, g):
    """(J^T g, dL/dm) of the direct sum: (J^T g)_i = sum_k T(d_ik)(m_i g_k - m_k g_i),
    dL/dm_i = -sum_k g_k . d_ik / s^3, T(d) = I/s^3 - 3 d d^T/s^5."""
    d = x[None, :, :] - x[:, None, :]
    s2 = (d * d).sum(2) + eps2
    s3 = 1.0 / (s2 * np.sqrt(s2))
    s5 = s3 / s2
    w = m[:, None, None] * g[None, :, :] - m[None, :, None] * g[:, None, :]
    dw = (d * w).sum(2)
    gx = (s3[:, :, None] * w - 3.0 * (s5 * dw)[:, :, None] * d).sum(1)
    gm = -((g[None, :, :] * d).sum(2) * s3).sum(1)
    return gx, gm


def accel(x, m, eps2, p):
    return direct_accel(x, m, eps2) + ext_accel(p, x)


def kdk_forward(x0, v0, m, eps2, dt, nsteps, pot=None):
    """States [(x, v, a)] for n = 0..nsteps of the KDK leapfrog."""
    p = pot_params(pot)
    x, v = np.array(x0, dtype=np.float64), np.array(v0, dtype=np.float64)
    a = accel(x, m, eps2, p)
    out = [(x.copy(), v.copy(), a.copy())]
    c = 0.5 * dt
    for _ in range(nsteps):
        vh = v + c * a
        x = x + dt * vh
        a = accel(x, m, eps2, p)
        v = vh + c * a
        out.append((x.copy(), v.copy(), a.copy()))
    return out


def kdk_adjoint(states, m, eps2, dt, gxN, gvN, pot=None):
    """Reverse sweep of kdk_forward: dL/d(x0, v0, m, dt) from dL/d(xN, vN)."""
    p = pot_params(pot)
    c = 0.5 * dt
    gx, gv = np.array(gxN, dtype=np.float64), np.array(gvN, dtype=np.float64)
    ga = np.zeros_like(gx)
    gm = np.zeros(m.shape[0])
    gdt = 0.0
    for n in range(len(states) - 2, -1, -1):
        x1, _, a1 = states[n + 1]
        _, v0, a0 = states[n]
        # v_{n+1} = v_h + c a_{n+1}
        ga = ga + c * gv
        gdt += 0.5 * (a1 * gv).sum()
        # a_{n+1} = A(x_{n+1})
        jx, jm = direct_adj(x1, m, eps2, ga)
        gx = gx + jx + ext_jac(p, x1, ga)
        gm = gm + jm
        # x_{n+1} = x_n + dt v_h ;  v_h = v_n + c a_n
        vh = v0 + c * a0
        gvh = gv + dt * gx
        gdt += (vh * gx).sum() + 0.5 * (a0 * gvh).sum()
        gv = gvh
        ga = c * gvh
    x0 = states[0][0]
    jx, jm = direct_adj(x0, m, eps2, ga)
    gx = gx + jx + ext_jac(p, x0, ga)
    gm = gm + jm
    return gx, gv, gm, gdt


# ---------------------------------------------------------------------------
# Frozen-tree force and its vector-Jacobian product
# ---------------------------------------------------------------------------
# The tree (cells, membership, opening decisions) is the oracle's, exported
# in pre-order (orc_tree_preorder: kind 0 = cell, 1 = body; skip = Next).
# frozen_accel recomputes every cell's mass, centre of mass and quadrupole
# from the body positions (hackCofM / hackQuad order: children in octant
# order, nbody/src/nbody_tree.c:70-121,308-424) and sums each body's fixed
# interaction list (nbGravity's terms, nbody/src/nbody_grav.c:53-99), so it
# is a smooth function of the positions and masses; frozen_vjp is its
# adjoint, written the way the GPU kernels compute it.

def tree_children(t):
    kind, skip = t["kind"], t["skip"]
    children = [[] for _ in range(len(kind))]
    for c in range(len(kind)):
        if kind[c] == 0:
            ch = c + 1
            while ch < skip[c]:
                children[c].append(ch)
                ch = skip[ch]
    return children


def walk_lists(t):
    """Per massive body: the nodes its reference walk accepts (self left out),
    replayed on the exported tree with the walk's arithmetic."""
    kind, ref, skip, pos, rc2 = t["kind"], t["ref"], t["skip"], t["pos"], t["rcrit2"]
    nn = len(kind)
    lists = {}
    for leaf in np.nonzero(kind == 1)[0]:
        b = int(ref[leaf])
        x, y, z = pos[leaf]
        lst, q = [], 0
        while q < nn:
            if kind[q] == 1:
                if ref[q] != b:
                    lst.append(q)
                q = int(skip[q])
            else:
                dx, dy, dz = pos[q, 0] - x, pos[q, 1] - y, pos[q, 2] - z
                if dz * dz + (dy * dy + dx * dx) >= rc2[q]:
                    lst.append(q)
                    q = int(skip[q])
                else:
                    q += 1
        lists[b] = np.array(lst, dtype=np.int64)
    return lists


def frozen_moments(X, m, t, children):
    kind, ref = t["kind"], t["ref"]
    nn = len(kind)
    cm, M, Q = np.zeros((nn, 3)), np.zeros(nn), np.zeros((nn, 6))
    for c in range(nn - 1, -1, -1):
        if kind[c] == 1:
            cm[c], M[c] = X[ref[c]], m[ref[c]]
            continue
        ch = children[c]
        M[c] = M[ch].sum()
        cm[c] = (M[ch, None] * cm[ch]).sum(0) / M[c]
        d = cm[ch] - cm[c]
        r2 = (d * d).sum(1)
        mm = M[ch]
        Q[c] = np.stack([mm * (3 * d[:, 0] * d[:, 0] - r2), mm * 3 * d[:, 0] * d[:, 1],
                         mm * 3 * d[:, 0] * d[:, 2], mm * (3 * d[:, 1] * d[:, 1] - r2),
                         mm * 3 * d[:, 1] * d[:, 2], mm * (3 * d[:, 2] * d[:, 2] - r2)], 1).sum(0) \
            + Q[ch].sum(0)
    return cm, M, Q


def _qmul(Q, v):
    return np.stack([Q[:, 0] * v[:, 0] + Q[:, 1] * v[:, 1] + Q[:, 2] * v[:, 2],
                     Q[:, 1] * v[:, 0] + Q[:, 3] * v[:, 1] + Q[:, 4] * v[:, 2],
                     Q[:, 2] * v[:, 0] + Q[:, 4] * v[:, 1] + Q[:, 5] * v[:, 2]], 1)


def frozen_accel(X, m, t, children, lists, eps2, quad=True):
    cm, M, Q = frozen_moments(X, m, t, children)
    cellq = (t["kind"] == 0) & quad
    a = np.zeros((X.shape[0], 3))
    for b, lst in lists.items():
        d = cm[lst] - X[b]
        s2 = (d * d).sum(1) + eps2
        s = np.sqrt(s2)
        acc = (M[lst] / s / s2)[:, None] * d
        qq = np.where(cellq[lst][:, None], Q[lst], 0.0)
        u = _qmul(qq, d)
        dqd = (u * d).sum(1)
        s5 = 1.0 / (s2 * s2 * s)
        acc += (2.5 * s5 * dqd / s2)[:, None] * d - s5[:, None] * u
        a[b] = acc.sum(0)
    return a


def frozen_vjp(X, m, t, children, lists, eps2, G, quad=True):
    """(dL/dX, dL/dm) of frozen_accel given dL/da = G."""
    kind, ref = t["kind"], t["ref"]
    nn = len(kind)
    cm, M, Q = frozen_moments(X, m, t, children)
    cellq = (kind == 0) & quad
    gx, gm = np.zeros_like(X), np.zeros(X.shape[0])
    gcm, gM, gQ = np.zeros((nn, 3)), np.zeros(nn), np.zeros((nn, 6))
    for b, lst in lists.items():
        g = G[b]
        d = cm[lst] - X[b]
        s2 = (d * d).sum(1) + eps2
        s3 = 1.0 / (s2 * np.sqrt(s2))
        s5, s7 = s3 / s2, s3 / s2 / s2
        s9 = s7 / s2
        dg = d @ g
        qq = np.where(cellq[lst][:, None], Q[lst], 0.0)
        u = _qmul(qq, d)
        v = _qmul(qq, np.broadcast_to(g, d.shape))
        dqd = (u * d).sum(1)
        ug = u @ g
        cg = M[lst] * s3 + 2.5 * dqd * s7
        cd = -3.0 * M[lst] * s5 * dg - 17.5 * dqd * s9 * dg + 5.0 * s7 * ug
        h = cg[:, None] * g + cd[:, None] * d + (5.0 * s7 * dg)[:, None] * u - s5[:, None] * v
        gx[b] -= h.sum(0)
        np.add.at(gcm, lst, h)
        np.add.at(gM, lst, s3 * dg)
        a2 = 2.5 * s7 * dg
        cQ = np.stack([a2 * d[:, 0] ** 2 - s5 * d[:, 0] * g[0],
                       2 * a2 * d[:, 0] * d[:, 1] - s5 * (d[:, 1] * g[0] + d[:, 0] * g[1]),
                       2 * a2 * d[:, 0] * d[:, 2] - s5 * (d[:, 2] * g[0] + d[:, 0] * g[2]),
                       a2 * d[:, 1] ** 2 - s5 * d[:, 1] * g[1],
                       2 * a2 * d[:, 1] * d[:, 2] - s5 * (d[:, 2] * g[1] + d[:, 1] * g[2]),
                       a2 * d[:, 2] ** 2 - s5 * d[:, 2] * g[2]], 1)
        np.add.at(gQ, lst, np.where(cellq[lst][:, None], cQ, 0.0))
    # cells hand their records to their children, roots first
    for c in range(nn):
        if kind[c] != 0:
            continue
        ch = np.array(children[c])
        d = cm[ch] - cm[c]
        w = M[ch] / M[c]
        gcm[ch] += w[:, None] * gcm[c]
        gM[ch] += gM[c] + (d @ gcm[c]) / M[c]
        if quad:
            G6 = gQ[c]
            mm = M[ch]
            gcm[ch, 0] += mm * (4 * d[:, 0] * G6[0] + 3 * d[:, 1] * G6[1] + 3 * d[:, 2] * G6[2]
                                - 2 * d[:, 0] * G6[3] - 2 * d[:, 0] * G6[5])
            gcm[ch, 1] += mm * (-2 * d[:, 1] * G6[0] + 3 * d[:, 0] * G6[1] + 4 * d[:, 1] * G6[3]
                                + 3 * d[:, 2] * G6[4] - 2 * d[:, 1] * G6[5])
            gcm[ch, 2] += mm * (-2 * d[:, 2] * G6[0] + 3 * d[:, 0] * G6[2] - 2 * d[:, 2] * G6[3]
                                + 3 * d[:, 1] * G6[4] + 4 * d[:, 2] * G6[5])
            r2 = (d * d).sum(1)
            gM[ch] += (G6[0] * (3 * d[:, 0] ** 2 - r2) + G6[1] * 3 * d[:, 0] * d[:, 1]
                       + G6[2] * 3 * d[:, 0] * d[:, 2] + G6[3] * (3 * d[:, 1] ** 2 - r2)
                       + G6[4] * 3 * d[:, 1] * d[:, 2] + G6[5] * (3 * d[:, 2] ** 2 - r2))
            cells = kind[ch] == 0
            gQ[ch[cells]] += G6
    leaves = np.nonzero(kind == 1)[0]
    gx[ref[leaves]] += gcm[leaves]
    gm[ref[leaves]] += gM[leaves]
    return gx, gm


# ---------------------------------------------------------------------------
# KDK over frozen trees: the map the tree adjoint differentiates
# ---------------------------------------------------------------------------
# Each state n of a tree run has its own tree and interaction lists (built
# at x_n by the run itself).  Holding them fixed while the inputs move
# gives a smooth map of (x0, v0, m, dt); its derivative is what
# nbhip_adjoint_backward computes for the tree criteria.  frozen[n] =
# (tree, children, lists) of state n, from the oracle's run.

def frozen_trees(oracle_state, nsteps):
    """Trees and lists of states 0..nsteps of an oracle run (force already
    evaluated at state 0); the oracle is left at state nsteps."""
    out = []
    for k in range(nsteps + 1):
        t = oracle_state.tree()
        out.append((t, tree_children(t), walk_lists(t)))
        if k < nsteps:
            st = oracle_state.step()
            if st != 0:
                raise RuntimeError(f"oracle step {k} returned status {st}")
    return out


def kdk_forward_frozen(x0, v0, m, eps2, dt, frozen, pot=None, quad=True):
    """States [(x, v, a)] of the KDK leapfrog whose force at state n is the
    frozen-list force of frozen[n] plus the external acceleration."""
    p = pot_params(pot)

    def force(x, n):
        t, ch, lists = frozen[n]
        return frozen_accel(x, m, t, ch, lists, eps2, quad) + ext_accel(p, x)

    x, v = np.array(x0, dtype=np.float64), np.array(v0, dtype=np.float64)
    a = force(x, 0)
    out = [(x.copy(), v.copy(), a.copy())]
    c = 0.5 * dt
    for n in range(1, len(frozen)):
        vh = v + c * a
        x = x + dt * vh
        a = force(x, n)
        v = vh + c * a
        out.append((x.copy(), v.copy(), a.copy()))
    return out


def kdk_adjoint_frozen(states, m, eps2, dt, frozen, gxN, gvN, pot=None, quad=True):
    """Reverse sweep of kdk_forward_frozen: dL/d(x0, v0, m, dt) from
    dL/d(xN, vN); the same recurrences as kdk_adjoint with frozen_vjp in
    place of the direct-sum Jacobian."""
    p = pot_params(pot)

    def vjp(x, n, g):
        t, ch, lists = frozen[n]
        jx, jm = frozen_vjp(x, m, t, ch, lists, eps2, g, quad)
        return jx + ext_jac(p, x, g), jm

    c = 0.5 * dt
    gx, gv = np.array(gxN, dtype=np.float64), np.array(gvN, dtype=np.float64)
    ga = np.zeros_like(gx)
    gm = np.zeros(m.shape[0])
    gdt = 0.0
    for n in range(len(states) - 2, -1, -1):
        x1, _, a1 = states[n + 1]
        _, v0, a0 = states[n]
        ga = ga + c * gv
        gdt += 0.5 * (a1 * gv).sum()
        jx, jm = vjp(x1, n + 1, ga)
        gx = gx + jx
        gm = gm + jm
        vh = v0 + c * a0
        gvh = gv + dt * gx
        gdt += (vh * gx).sum() + 0.5 * (a0 * gvh).sum()
        gv = gvh
        ga = c * gvh
    jx, jm = vjp(states[0][0], 0, ga)
    return gx + jx, gv, gm + jm, gdt
