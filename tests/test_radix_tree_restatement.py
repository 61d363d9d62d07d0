"""The tree build's derivations, restated in Python and checked against
brute force on the CPU (the GPU tree tests check the kernels themselves, bit
for bit, against the oracle's tree).

* Karras's binary radix tree (T. Karras, HPG 2012: per internal node, binary
  searches over the keys for its range and split) equals the minimum
  Cartesian tree of the adjacent common prefixes: the node split at k covers
  L+1 .. R, L and R the nearest smaller prefixes to its left and right, and
  hangs below whichever of the two is larger.  Keys carry a simulation id
  above the Morton bits and the body index as the tie-break of equal keys.
* The nearest-smaller search through a min-hierarchy over the prefixes (16
  per entry and level, levels starting on multiples of 16, up through the
  siblings of the chunk passed and down into the first chunk holding a
  smaller prefix) finds the brute-force answer, chunk edges included.
* The octree parent of a cell or a body's leaf is the level-P cell holding a
  given body, whose first body follows the nearest prefix left of it below
  level P: the same index as a walk up the binary tree to the nearest
  ancestor that owns an octree cell.
"""
import numpy as np
import pytest

KEY_BITS = 20          # Morton bits of the restated keys
INF = 2 ** 31 - 1


def clz32(x):
    return 32 - int(x).bit_length()


def make_keys(rng, m, sb, nb):
    keys = [int(x) for x in rng.integers(0, 1 << nb, m)]
    sims = [int(x) for x in rng.integers(0, (1 << sb) if sb else 1, m)]
    pairs = sorted(zip(sims, keys))
    return [p[1] for p in pairs], [p[0] for p in pairs]


def prefix_fn(keys, sims, sb):
    """Common prefix of sorted bodies i and j: simulation bits, then key
    bits, then the index tie-break; -1 outside [0, M)."""
    m = len(keys)

    def delta(i, j):
        if j < 0 or j >= m:
            return -1
        if sims[i] != sims[j]:
            return clz32(sims[i] ^ sims[j]) - (32 - sb)
        x = keys[i] ^ keys[j]
        d = KEY_BITS - x.bit_length() if x else KEY_BITS
        if d == KEY_BITS:
            d = KEY_BITS + clz32(i ^ j)
        return sb + d
    return delta


def karras(delta, m):
    """Karras's construction: (lo, hi, prefix) per internal node and the
    parent of every node (internal 0..m-2, leaf a at m-1+a)."""
    lo_, hi_, dn_, par = [None] * (m - 1), [None] * (m - 1), [None] * (m - 1), [None] * (2 * m - 1)
    for i in range(m - 1):
        d = 1 if delta(i, i + 1) > delta(i, i - 1) else -1
        dmin = delta(i, i - d)
        lmax = 2
        while delta(i, i + lmax * d) > dmin:
            lmax <<= 1
        ln, t = 0, lmax >> 1
        while t >= 1:
            if delta(i, i + (ln + t) * d) > dmin:
                ln += t
            t >>= 1
        j = i + ln * d
        dnode = delta(i, j)
        s, div = 0, 2
        while True:
            t = (ln + div - 1) // div
            if delta(i, i + (s + t) * d) > dnode:
                s += t
            if t <= 1:
                break
            div <<= 1
        g = i + s * d + min(d, 0)
        lo, hi = min(i, j), max(i, j)
        lo_[i], hi_[i], dn_[i] = lo, hi, dnode
        par[(m - 1) + g if g == lo else g] = i
        par[(m - 1) + g + 1 if g + 1 == hi else g + 1] = i
    par[0] = -1
    return lo_, hi_, dn_, par


def cartesian(dadj, m):
    """The same outputs from the nearest smaller prefixes (brute force)."""
    lo_, hi_, dn_, par = [None] * (m - 1), [None] * (m - 1), [None] * (m - 1), [None] * (2 * m - 1)
    for k in range(m - 1):
        v = dadj[k]
        r = next(p for p in range(k + 1, m) if dadj[p] < v)
        left = next((p for p in range(k - 1, -1, -1) if dadj[p] < v), -1)
        lo, hi = left + 1, r
        vl = dadj[left] if left >= 0 else -1
        vr = dadj[r] if r < m - 1 else -1
        node = 0 if (vl < 0 and vr < 0) else (hi if vr > vl else lo)
        lo_[node], hi_[node], dn_[node] = lo, hi, v
        par[(m - 1) + k if k == lo else k] = node
        par[(m - 1) + k + 1 if k + 1 == hi else k + 1] = node
    par[0] = -1
    return lo_, hi_, dn_, par


@pytest.mark.parametrize("seed", range(6))
def test_karras_tree_is_the_cartesian_tree_of_the_prefixes(seed):
    rng = np.random.default_rng(seed)
    for _ in range(40):
        m = int(rng.integers(2, 160))
        sb = int(rng.integers(0, 4))
        nb = int(rng.integers(1, KEY_BITS + 1))        # few bits: many equal keys
        keys, sims = make_keys(rng, m, sb, nb)
        delta = prefix_fn(keys, sims, sb)
        dadj = [delta(a, a + 1) for a in range(m)]
        assert karras(delta, m) == cartesian(dadj, m)


# --- the min-hierarchy search, as the kernels index it ------------------------

def hierarchy(dadj, n):
    """dadj's allocation: n prefix slots, then levels L >= 1 (minimum of 16
    entries of level L-1, +inf past the end), each starting on a multiple of
    16; returns the buffer and per level (size, offset)."""
    m = len(dadj)
    sizes, offs = [m], [0]
    at = (n + 15) & ~15
    while sizes[-1] > 16:
        s = (sizes[-1] + 15) // 16
        sizes.append(s)
        offs.append(at)
        at += (s + 15) & ~15
    buf = np.full(at + 16, -7, dtype=np.int64)         # unwritten slots: below any prefix
    buf[:m] = dadj
    for lv in range(1, len(sizes)):
        prev = np.full(sizes[lv] * 16, INF, dtype=np.int64)
        prev[:sizes[lv - 1]] = buf[offs[lv - 1]:offs[lv - 1] + sizes[lv - 1]]
        buf[offs[lv]:offs[lv] + sizes[lv]] = prev.reshape(-1, 16).min(1)
    return buf, sizes, offs


def below(buf, sizes, offs, lv, base, v):
    """Bit t: entry base + t of level lv is below v (the kernel's 64-byte
    chunk, bounded by the level's size)."""
    assert base % 16 == 0 and (offs[lv] + base) % 16 == 0
    assert offs[lv] + base + 16 <= len(buf)             # the chunk stays in the allocation
    m = 0
    for t in range(16):
        if int(buf[offs[lv] + base + t]) < v:
            m |= 1 << t
    lim = sizes[lv] - base
    return m if lim >= 16 else m & ((1 << max(lim, 0)) - 1)


def nearest_smaller(buf, sizes, offs, i, v, right):
    top = len(sizes) - 1
    c, lv = i, 0
    while True:
        base, t = c & ~15, c & 15
        m = below(buf, sizes, offs, lv, base, v)
        m &= (~((2 << t) - 1) & 0xFFFF) if right else ((1 << t) - 1)
        if m:
            c = base + ((m & -m).bit_length() - 1 if right else m.bit_length() - 1)
            break
        if lv >= top:
            return -2 if right else -1
        c >>= 4
        lv += 1
    while lv > 0:
        m = below(buf, sizes, offs, lv - 1, c << 4, v)
        assert m, "a chunk whose minimum is below v holds an entry below v"
        c = (c << 4) + ((m & -m).bit_length() - 1 if right else m.bit_length() - 1)
        lv -= 1
    return c


@pytest.mark.parametrize("m", [2, 3, 15, 16, 17, 255, 256, 257, 4095, 4097, 5000])
def test_hierarchy_search_finds_the_nearest_smaller_prefix(m):
    rng = np.random.default_rng(m)
    n = m + int(rng.integers(0, 40))
    dadj = rng.integers(0, 60, m).astype(np.int64)
    dadj[m - 1] = -1                                    # the last body's prefix: none
    buf, sizes, offs = hierarchy(dadj, n)
    ks = range(m - 1) if m <= 300 else rng.integers(0, m - 1, 300)
    for k in ks:
        k, v = int(k), int(dadj[k])
        r = next(p for p in range(k + 1, m) if dadj[p] < v)
        left = next((p for p in range(k - 1, -1, -1) if dadj[p] < v), -1)
        assert nearest_smaller(buf, sizes, offs, k, v, True) == r
        assert nearest_smaller(buf, sizes, offs, k, v, False) == left


# --- octree parents -----------------------------------------------------------

def levels(bits, sb):
    return -1 if bits < sb else min(bits - sb, 189) // 3


@pytest.mark.parametrize("seed", range(4))
def test_octree_parents_from_the_nearest_smaller_prefix(seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(40):
        m = int(rng.integers(2, 160))
        sb = int(rng.integers(0, 4))
        nb = int(rng.integers(3, KEY_BITS + 1))
        keys, sims = make_keys(rng, m, sb, nb)
        delta = prefix_fn(keys, sims, sb)
        dadj = [delta(a, a + 1) for a in range(m)]
        lo_, hi_, dn_, par = karras(delta, m)
        first_of_sim = [a == 0 or sims[a - 1] != sims[a] for a in range(m)]
        cells = [max(levels(dadj[a], sb), 0) + 1 if first_of_sim[a]
                 else max(0, levels(dadj[a], sb) - levels(dadj[a - 1], sb)) for a in range(m)]
        cincl = list(np.cumsum(cells))
        excl = lambda a: 0 if a == 0 else cincl[a - 1]
        first_level = lambda a: 0 if a == 0 else levels(dadj[a - 1], sb) + 1
        parent_levels = lambda k: -1 if k == 0 else levels(dn_[par[k]], sb)

        def owner_pre(k):           # up the binary tree to a node that owns a cell
            while k > 0 and levels(dn_[k], sb) <= parent_levels(k):
                k = par[k]
            lo = lo_[k]
            return lo + excl(lo) + (levels(dn_[k], sb) - first_level(lo))

        def cell_at(i, p):          # the level-p cell holding body i
            a = next((q for q in range(i - 1, -1, -1) if dadj[q] < sb + 3 * p), -1) + 1
            return a + excl(a) + (p - first_level(a))

        for s in range(m):
            single = first_of_sim[s] and (s + 1 == m or sims[s + 1] != sims[s])
            if single:
                continue
            p = max(levels(dadj[s - 1] if s > 0 else -1, sb), levels(dadj[s], sb))
            assert cell_at(s, p) == owner_pre(par[(m - 1) + s])
        for k in range(m - 1):
            lk, lp = levels(dn_[k], sb), parent_levels(k)
            if lk <= lp or lp < 0:
                continue
            lo, hi = lo_[k], hi_[k]
            vl = dadj[lo - 1] if lo > 0 else -1
            vr = dadj[hi] if hi < m - 1 else -1
            assert levels(max(vl, vr), sb) == lp
            assert cell_at(lo, lp) == owner_pre(par[k])
