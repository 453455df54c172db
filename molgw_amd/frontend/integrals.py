"""Gaussian integrals by the McMurchie-Davidson (Hermite Gaussian) scheme,
written from the mathematics (Helgaker, Jorgensen, Olsen, Molecular
Electronic-Structure Theory, ch. 9): overlap, kinetic energy, nuclear
attraction, and the two- and three-centre Coulomb integrals of the RI basis
(the integrals the reference takes from libcint: src/m_eri_calculate.f90,
src/m_hamiltonian_onebody.f90).

Shell-level routines return the integrals over the Cartesian components of
the contracted shells (x^i y^j z^k exp(-a r^2), weights Shell.coefs); the
molecule-level routines apply each shell's transform to the final functions.
Every routine is vectorised over the primitives of the shells involved.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import gammainc, gammaln

from .basis import BasisSet, Shell, cart_components


# ---------------------------------------------------------------- Boys function
def boys(nmax: int, t: np.ndarray) -> np.ndarray:
    """F_n(t) = int_0^1 u^(2n) exp(-t u^2) du for n = 0..nmax; shape (nmax+1, *t.shape).
    Small t: the Taylor series; otherwise F_nmax from the regularised lower
    incomplete gamma function, then the stable downward recursion
    F_n = (2 t F_(n+1) + exp(-t)) / (2n + 1)."""
    t = np.asarray(t, dtype=np.float64)
    out = np.empty((nmax + 1,) + t.shape)
    small = t < 1.0
    ts = t[small]
    tl = t[~small]
    if ts.size:
        for n in range(nmax + 1):
            term = np.ones_like(ts)
            acc = term / (2 * n + 1)
            for k in range(1, 40):
                term = term * (-ts) / k
                acc = acc + term / (2 * n + 2 * k + 1)
            out[n][small] = acc
    if tl.size:
        a = nmax + 0.5
        f = np.exp(gammaln(a)) * gammainc(a, tl) / (2.0 * tl ** a)
        out[nmax][~small] = f
        e = np.exp(-tl)
        for n in range(nmax - 1, -1, -1):
            f = (2.0 * tl * f + e) / (2 * n + 1)
            out[n][~small] = f
    return out


# ---------------------------------------------------------------- Hermite expansion
def hermite_e(la: int, lb: int, a, b, xab):
    """E[i, j, t] of the product of two 1-D Gaussians (exponents a, b; centre
    separation xab = A - B), i <= la, j <= lb, t <= i + j; trailing dims are
    those of the broadcast a, b, xab."""
    a, b, xab = np.broadcast_arrays(*(np.asarray(v, dtype=np.float64) for v in (a, b, xab)))
    p = a + b
    mu = a * b / p
    xpa = -b * xab / p
    xpb = a * xab / p
    e = np.zeros((la + 1, lb + 1, la + lb + 2) + a.shape)
    e[0, 0, 0] = np.exp(-mu * xab * xab)
    inv2p = 0.5 / p
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            if i > 0:
                src, x = e[i - 1, j], xpa
            else:
                src, x = e[i, j - 1], xpb
            for t in range(i + j + 1):
                v = x * src[t]
                if t > 0:
                    v = v + inv2p * src[t - 1]
                if t + 1 <= i + j - 1:
                    v = v + (t + 1) * src[t + 1]
                e[i, j, t] = v
    return e[:, :, :la + lb + 1]


def hermite_r(lmax: int, alpha, pc):
    """R[t, u, v] (t + u + v <= lmax) of the Hermite Coulomb integrals for
    exponent alpha and separation pc (..., 3); trailing dims broadcast."""
    alpha = np.asarray(alpha, dtype=np.float64)
    pc = np.asarray(pc, dtype=np.float64)
    x, y, z = pc[..., 0], pc[..., 1], pc[..., 2]
    shape = np.broadcast_shapes(alpha.shape, x.shape)
    t2 = alpha * (x * x + y * y + z * z)
    f = boys(lmax, np.broadcast_to(t2, shape))
    n1 = lmax + 1
    # level n holds R^n[t, u, v] for t + u + v <= lmax - n
    cur = np.zeros((n1, n1, n1) + shape)
    cur[0, 0, 0] = (-2.0 * alpha) ** lmax * f[lmax]
    for n in range(lmax - 1, -1, -1):
        nxt = np.zeros_like(cur)
        nxt[0, 0, 0] = (-2.0 * alpha) ** n * f[n]
        m = lmax - n
        for t in range(m + 1):
            for u in range(m + 1 - t):
                for v in range(m + 1 - t - u):
                    if t == u == v == 0:
                        continue
                    if t > 0:
                        r = x * cur[t - 1, u, v]
                        if t > 1:
                            r = r + (t - 1) * cur[t - 2, u, v]
                    elif u > 0:
                        r = y * cur[t, u - 1, v]
                        if u > 1:
                            r = r + (u - 1) * cur[t, u - 2, v]
                    else:
                        r = z * cur[t, u, v - 1]
                        if v > 1:
                            r = r + (v - 1) * cur[t, u, v - 2]
                    nxt[t, u, v] = r
        cur = nxt
    return cur


# ---------------------------------------------------------------- shell pairs
def _pair(sa: Shell, sb: Shell, lb_extra: int = 0):
    """Hermite data of a shell pair over its primitive pairs (na, nb)."""
    a = sa.exps[:, None]
    b = sb.exps[None, :]
    p = a + b
    ab = sa.center - sb.center
    ex = [hermite_e(sa.l, sb.l + lb_extra, a, b, ab[d]) for d in range(3)]
    pcen = (a[..., None] * sa.center + b[..., None] * sb.center) / p[..., None]
    w = sa.coefs[:, None] * sb.coefs[None, :]
    return p, pcen, ex, w


def _hermite_pair(sa: Shell, sb: Shell, ex):
    """H[ca, cb, t, u, v, na, nb] = Ex[ia, ib, t] Ey[ja, jb, u] Ez[ka, kb, v]."""
    L = sa.l + sb.l
    ca, cb = cart_components(sa.l), cart_components(sb.l)
    h = np.zeros((len(ca), len(cb), L + 1, L + 1, L + 1) + ex[0].shape[3:])
    for i, (ax, ay, az) in enumerate(ca):
        for j, (bx, by, bz) in enumerate(cb):
            h[i, j] = (ex[0][ax, bx][:, None, None] * ex[1][ay, by][None, :, None] *
                       ex[2][az, bz][None, None, :])
    return h


def _hermite_single(sc: Shell):
    """G[cc, t, u, v, nc] of a single shell (Hermite expansion of one Gaussian)."""
    q = sc.exps
    ex = hermite_e(sc.l, 0, q, np.zeros_like(q), np.zeros_like(q))   # partner of exponent 0
    cc = cart_components(sc.l)
    L = sc.l
    g = np.zeros((len(cc), L + 1, L + 1, L + 1, len(q)))
    for i, (cx, cy, cz) in enumerate(cc):
        g[i] = ex[cx, 0][:, None, None] * ex[cy, 0][None, :, None] * ex[cz, 0][None, None, :]
    return g


def shell_overlap_cart(sa: Shell, sb: Shell) -> np.ndarray:
    p, _, ex, w = _pair(sa, sb)
    pref = (math.pi / p) ** 1.5 * w
    ca, cb = cart_components(sa.l), cart_components(sb.l)
    s = np.zeros((len(ca), len(cb)))
    for i, (ax, ay, az) in enumerate(ca):
        for j, (bx, by, bz) in enumerate(cb):
            s[i, j] = np.sum(ex[0][ax, bx, 0] * ex[1][ay, by, 0] * ex[2][az, bz, 0] * pref)
    return s


def shell_kinetic_cart(sa: Shell, sb: Shell) -> np.ndarray:
    """-1/2 <a|nabla^2|b>: per direction T_ij = -2 b^2 S_(i,j+2) + b (2j+1) S_ij - j(j-1)/2 S_(i,j-2)."""
    p, _, ex, w = _pair(sa, sb, lb_extra=2)
    b = sb.exps[None, :]
    s1 = [e[:, :, 0] * np.sqrt(math.pi / p) for e in ex]      # 1-D overlaps [i, j, na, nb]

    def t1(d, i, j):
        v = -2.0 * b * b * s1[d][i, j + 2] + b * (2 * j + 1) * s1[d][i, j]
        if j >= 2:
            v = v - 0.5 * j * (j - 1) * s1[d][i, j - 2]
        return v

    ca, cb = cart_components(sa.l), cart_components(sb.l)
    out = np.zeros((len(ca), len(cb)))
    for i, a_ in enumerate(ca):
        for j, b_ in enumerate(cb):
            sx, sy, sz = (s1[d][a_[d], b_[d]] for d in range(3))
            tx, ty, tz = (t1(d, a_[d], b_[d]) for d in range(3))
            out[i, j] = np.sum((tx * sy * sz + sx * ty * sz + sx * sy * tz) * w)
    return out


def shell_nuclear_cart(sa: Shell, sb: Shell, charges, coords) -> np.ndarray:
    """sum_C -Z_C <a| 1/|r - C| |b>."""
    p, pcen, ex, w = _pair(sa, sb)
    h = _hermite_pair(sa, sb, ex)
    L = sa.l + sb.l
    out = np.zeros(h.shape[:2])
    for z, c in zip(charges, coords):
        r = hermite_r(L, p, pcen - np.asarray(c))                       # [t, u, v, na, nb]
        out -= z * np.einsum('ijtuvab,tuvab,ab->ij', h, r, 2.0 * math.pi / p * w, optimize=True)
    return out


def _ket_signs(l: int) -> np.ndarray:
    n = l + 1
    t = np.arange(n)
    return (-1.0) ** (t[:, None, None] + t[None, :, None] + t[None, None, :])


def _gather_r(r, n1: int, m1: int):
    """R[t + T, u + U, v + V] as [t, u, v, T, U, V, ...]."""
    it = np.arange(n1)[:, None] + np.arange(m1)[None, :]
    return r[it[:, None, None, :, None, None], it[None, :, None, None, :, None], it[None, None, :, None, None, :]]


def shell_eri3_group(sa: Shell, sb: Shell, group, pair=None) -> list:
    """(ab|c) over the Cartesian components for a group of auxiliary shells c
    of one angular momentum, their primitives batched together; returns a
    list of (ca, cb, cc) arrays, one per shell of the group."""
    p, pcen, ex, w = pair if pair is not None else _pair(sa, sb)
    hab = _hermite_pair(sa, sb, ex)                                     # [ca, cb, t, u, v, na, nb]
    lc = group[0].l
    q = np.concatenate([sc.exps for sc in group])
    coef = np.concatenate([sc.coefs for sc in group])
    cen = np.concatenate([np.broadcast_to(sc.center, (len(sc.exps), 3)) for sc in group])
    seg = np.cumsum([0] + [len(sc.exps) for sc in group])
    one = Shell(lc, np.zeros(3), q, coef, -1)
    g = _hermite_single(one) * _ket_signs(lc)[None, :, :, :, None]     # [cc, T, U, V, nc]
    Lab = sa.l + sb.l
    pp = p[:, :, None]
    qq = q[None, None, :]
    alpha = pp * qq / (pp + qq)
    r = hermite_r(Lab + lc, alpha, pcen[:, :, None, :] - cen[None, None, :, :])   # [t', u', v', na, nb, nc]
    pref = 2.0 * math.pi ** 2.5 / (pp * qq * np.sqrt(pp + qq)) * w[:, :, None] * coef[None, None, :]
    rg = _gather_r(r, Lab + 1, lc + 1) * pref                           # [t, u, v, T, U, V, na, nb, nc]
    x = np.einsum('ijtuvab,tuvTUVabc->ijTUVc', hab, rg, optimize=True)
    y = np.einsum('ijTUVc,kTUVc->ijkc', x, g, optimize=True)
    return [y[..., seg[k]:seg[k + 1]].sum(axis=-1) for k in range(len(group))]


def shell_eri2_cart(sc: Shell, sd: Shell) -> np.ndarray:
    """(c|d) over the Cartesian components, shape (cc, cd)."""
    gc = _hermite_single(sc)                                            # [cc, t, u, v, nc]
    gd = _hermite_single(sd) * _ket_signs(sd.l)[None, :, :, :, None]
    p = sc.exps[:, None]
    q = sd.exps[None, :]
    alpha = p * q / (p + q)
    r = hermite_r(sc.l + sd.l, alpha, np.broadcast_to(sc.center - sd.center, p.shape[:1] + q.shape[1:] + (3,)))
    pref = 2.0 * math.pi ** 2.5 / (p * q * np.sqrt(p + q)) * sc.coefs[:, None] * sd.coefs[None, :]
    rg = _gather_r(r, sc.l + 1, sd.l + 1)
    x = np.einsum('ktuvc,tuvTUVcd,cd->kTUVd', gc, rg, pref, optimize=True)
    return np.einsum('kTUVd,lTUVd->kl', x, gd, optimize=True)


# ---------------------------------------------------------------- molecule level
def _blocks(bs: BasisSet):
    return [(sh, int(bs.offsets[k]), int(bs.offsets[k + 1])) for k, sh in enumerate(bs.shells)]


def one_electron(bs: BasisSet, charges, coords):
    """(S, T, V) over the final functions."""
    n = bs.nbf
    s, t, v = (np.zeros((n, n)) for _ in range(3))
    blk = _blocks(bs)
    for i, (sa, a0, a1) in enumerate(blk):
        for k, (sb, b0, b1) in enumerate(blk[:i + 1]):
            ta, tb = sa.transform, sb.transform
            for mat, fn in ((s, shell_overlap_cart), (t, shell_kinetic_cart), (v, None)):
                m = ta @ (fn(sa, sb) if fn else shell_nuclear_cart(sa, sb, charges, coords)) @ tb.T
                if k == i:
                    m = 0.5 * (m + m.T)       # exactly symmetric diagonal blocks
                mat[a0:a1, b0:b1] = m
                mat[b0:b1, a0:a1] = m.T
    return s, t, v


def eri_2center(aux: BasisSet) -> np.ndarray:
    """(P|Q) over the auxiliary functions."""
    n = aux.nbf
    out = np.zeros((n, n))
    blk = _blocks(aux)
    for i, (sc, c0, c1) in enumerate(blk):
        for k, (sd, d0, d1) in enumerate(blk[:i + 1]):
            m = sc.transform @ shell_eri2_cart(sc, sd) @ sd.transform.T
            if k == i:
                m = 0.5 * (m + m.T)
            out[c0:c1, d0:d1] = m
            out[d0:d1, c0:c1] = m.T
    return out


def eri_3center(bs: BasisSet, aux: BasisSet) -> np.ndarray:
    """(mu nu|P) over the final functions, dense (nbf, nbf, naux). The
    auxiliary shells of one angular momentum are batched per shell pair."""
    n, na = bs.nbf, aux.nbf
    out = np.zeros((n, n, na))
    blk, ablk = _blocks(bs), _blocks(aux)
    groups = {}
    for sc, c0, c1 in ablk:
        groups.setdefault(sc.l, []).append((sc, c0, c1))
    for i, (sa, a0, a1) in enumerate(blk):
        for k, (sb, b0, b1) in enumerate(blk[:i + 1]):
            pair = _pair(sa, sb)
            for lc, members in groups.items():
                res = shell_eri3_group(sa, sb, [m[0] for m in members], pair)
                for (sc, c0, c1), m in zip(members, res):
                    m = np.einsum('ia,jb,kc,abc->ijk', sa.transform, sb.transform, sc.transform, m, optimize=True)
                    if k == i:
                        m = 0.5 * (m + m.transpose(1, 0, 2))
                    out[a0:a1, b0:b1, c0:c1] = m
                    out[b0:b1, a0:a1, c0:c1] = m.transpose(1, 0, 2)
    return out


def nuclear_repulsion(charges, coords) -> float:
    e = 0.0
    for i in range(len(charges)):
        for j in range(i):
            e += charges[i] * charges[j] / np.linalg.norm(np.asarray(coords[i]) - np.asarray(coords[j]))
    return e
