"""Independent textbook ERI (McMurchie-Davidson, Helgaker/Jorgensen/Olsen,
"Molecular Electronic-Structure Theory", sec. 9.5 and 9.9) in numpy/scipy.

Used only to pin the oracle's Rys-quadrature ERIs: a different algorithm
(Hermite Gaussians + Boys function from scipy's incomplete gamma), so a
dropped term, wrong sign or transposed index in either one fails the test.
"""
import math

import numpy as np
from scipy.special import gammainc, gamma

FAC_SP = {0: 0.282094791773878143, 1: 0.488602511902919921}


def boys(nmax, x):
    out = np.empty(nmax + 1)
    for m in range(nmax + 1):
        if x < 1e-8:
            out[m] = 1.0 / (2 * m + 1) - x / (2 * m + 3)
        else:
            a = m + 0.5
            out[m] = gamma(a) * gammainc(a, x) / (2.0 * x ** a)
    return out


def hermite_E(i, j, a, b, XAB):
    """E[i][j][t] for one axis, including exp(-mu X_AB^2)."""
    p = a + b
    mu = a * b / p
    XPA = -b / p * XAB
    XPB = a / p * XAB
    E = np.zeros((i + 1, j + 1, i + j + 2))
    E[0, 0, 0] = math.exp(-mu * XAB * XAB)
    for ii in range(i + 1):
        for jj in range(j + 1):
            if ii == 0 and jj == 0:
                continue
            for t in range(ii + jj + 1):
                if ii > 0:
                    v = XPA * E[ii - 1, jj, t] + (t + 1) * E[ii - 1, jj, t + 1]
                    if t > 0:
                        v += E[ii - 1, jj, t - 1] / (2 * p)
                else:
                    v = XPB * E[ii, jj - 1, t] + (t + 1) * E[ii, jj - 1, t + 1]
                    if t > 0:
                        v += E[ii, jj - 1, t - 1] / (2 * p)
                E[ii, jj, t] = v
    return E


def hermite_R(L, alpha, PQ):
    """R[t][u][v] = R^0_{tuv}(alpha, PQ) for t+u+v <= L."""
    X, Y, Z = PQ
    F = boys(L, alpha * (X * X + Y * Y + Z * Z))
    Rn = np.zeros((L + 1, L + 1, L + 1, L + 1))
    for n in range(L + 1):
        Rn[n, 0, 0, 0] = (-2 * alpha) ** n * F[n]
    for tot in range(1, L + 1):
        for n in range(L - tot + 1):
            for t in range(tot + 1):
                for u in range(tot - t + 1):
                    v = tot - t - u
                    if t > 0:
                        val = X * Rn[n + 1, t - 1, u, v] + ((t - 1) * Rn[n + 1, t - 2, u, v] if t > 1 else 0)
                    elif u > 0:
                        val = Y * Rn[n + 1, t, u - 1, v] + ((u - 1) * Rn[n + 1, t, u - 2, v] if u > 1 else 0)
                    else:
                        val = Z * Rn[n + 1, t, u, v - 1] + ((v - 1) * Rn[n + 1, t, u, v - 2] if v > 1 else 0)
                    Rn[n, t, u, v] = val
    return Rn[0]


def cart(l):
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def shell(mol, s, ctr=0):
    """(l, centre, exponents, coefficients of contraction ctr times the
    l = 0, 1 normalisation factor)."""
    b = mol._bas[s]
    a = mol._atm[b[0]]
    R = mol._env[a[1]:a[1] + 3]
    l, nprim = int(b[1]), int(b[2])
    exps = mol._env[b[5]:b[5] + nprim]
    coefs = mol._env[b[6] + ctr * nprim:b[6] + (ctr + 1) * nprim] * FAC_SP.get(l, 1.0)
    return l, R, exps, coefs


def eri_block(mol, shls):
    """(ij|kl) block, shape (di, dj, dk, dl) with d = nf * nctr and the
    component index fastest within each contraction (I = f + nf * c), the
    layout of a generally contracted shell in cint2e_cart."""
    nct = [int(mol._bas[s][3]) for s in shls]
    nfs = [len(cart(int(mol._bas[s][1]))) for s in shls]
    out = np.zeros([nf * nc for nf, nc in zip(nfs, nct)])
    for c0 in range(nct[0]):
        for c1 in range(nct[1]):
            for c2 in range(nct[2]):
                for c3 in range(nct[3]):
                    blk = _eri_block_ctr(mol, shls, (c0, c1, c2, c3))
                    out[c0 * nfs[0]:(c0 + 1) * nfs[0], c1 * nfs[1]:(c1 + 1) * nfs[1],
                        c2 * nfs[2]:(c2 + 1) * nfs[2], c3 * nfs[3]:(c3 + 1) * nfs[3]] = blk
    return out


def _eri_block_ctr(mol, shls, ctr):
    """(ij|kl) block of one contraction of each shell, shape (nfi, nfj, nfk, nfl)."""
    sh = [shell(mol, s, c) for s, c in zip(shls, ctr)]
    (la, A, ea, ca), (lb, B, eb, cb), (lc, C, ec, cc), (ld, D, ed, cd) = sh
    fa, fb, fc, fd = cart(la), cart(lb), cart(lc), cart(ld)
    out = np.zeros((len(fa), len(fb), len(fc), len(fd)))
    L = la + lb + lc + ld
    for a, c1 in zip(ea, ca):
        for b, c2 in zip(eb, cb):
            p = a + b
            P = (a * A + b * B) / p
            Eab = [hermite_E(la, lb, a, b, A[x] - B[x]) for x in range(3)]
            for c, c3 in zip(ec, cc):
                for d, c4 in zip(ed, cd):
                    q = c + d
                    Q = (c * C + d * D) / q
                    Ecd = [hermite_E(lc, ld, c, d, C[x] - D[x]) for x in range(3)]
                    alpha = p * q / (p + q)
                    R = hermite_R(L, alpha, P - Q)
                    pref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * c1 * c2 * c3 * c4
                    for ia, (ax, ay, az) in enumerate(fa):
                        for ib, (bx, by, bz) in enumerate(fb):
                            for ic, (cx, cy, cz) in enumerate(fc):
                                for id_, (dx, dy, dz) in enumerate(fd):
                                    s = 0.0
                                    for t in range(ax + bx + 1):
                                        for u in range(ay + by + 1):
                                            for v in range(az + bz + 1):
                                                e1 = Eab[0][ax, bx, t] * Eab[1][ay, by, u] * Eab[2][az, bz, v]
                                                if e1 == 0.0:
                                                    continue
                                                for tau in range(cx + dx + 1):
                                                    for nu in range(cy + dy + 1):
                                                        for phi in range(cz + dz + 1):
                                                            e2 = Ecd[0][cx, dx, tau] * Ecd[1][cy, dy, nu] * Ecd[2][cz, dz, phi]
                                                            sgn = -1.0 if (tau + nu + phi) % 2 else 1.0
                                                            s += e1 * sgn * e2 * R[t + tau, u + nu, v + phi]
                                    out[ia, ib, ic, id_] += pref * s
    return out


def _ovlp_1d(i, j, a, b, XAB):
    """1D overlap of x_A^i x_B^j exp(-a x_A^2 - b x_B^2), E^{ij}_0 sqrt(pi/p)."""
    if i < 0 or j < 0:
        return 0.0
    return hermite_E(i, j, a, b, XAB)[i, j, 0] * math.sqrt(math.pi / (a + b))


def int1e_block(mol, i, j, typ):
    """One-electron block (nfi*nctr_i, nfj*nctr_j), I = f + nf * c, by
    McMurchie-Davidson: typ 0 overlap, 1 kinetic -1/2 <i|lap|j>, 2 nuclear
    sum_C -Z_C <i|1/r_C|j>."""
    ni, nj = int(mol._bas[i][3]), int(mol._bas[j][3])
    fi = len(cart(int(mol._bas[i][1])))
    fj = len(cart(int(mol._bas[j][1])))
    out = np.zeros((fi * ni, fj * nj))
    for ci in range(ni):
        for cj in range(nj):
            out[ci * fi:(ci + 1) * fi, cj * fj:(cj + 1) * fj] = _int1e_block_ctr(mol, i, j, typ, ci, cj)
    return out


def _int1e_block_ctr(mol, i, j, typ, ci, cj):
    li, A, ea, ca = shell(mol, i, ci)
    lj, B, eb, cb = shell(mol, j, cj)
    fi, fj = cart(li), cart(lj)
    out = np.zeros((len(fi), len(fj)))
    atoms = [(abs(int(a[0])), mol._env[a[1]:a[1] + 3]) for a in mol._atm]
    for a, c1 in zip(ea, ca):
        for b, c2 in zip(eb, cb):
            p = a + b
            P = (a * A + b * B) / p
            E = [hermite_E(li, lj + 2, a, b, A[x] - B[x]) for x in range(3)]
            for ia, pa in enumerate(fi):
                for jb, pb in enumerate(fj):
                    if typ == 0:
                        v = np.prod([_ovlp_1d(pa[x], pb[x], a, b, A[x] - B[x]) for x in range(3)])
                    elif typ == 1:
                        S = [_ovlp_1d(pa[x], pb[x], a, b, A[x] - B[x]) for x in range(3)]
                        v = 0.0
                        for x in range(3):
                            jx = pb[x]
                            tx = (-0.5 * jx * (jx - 1) * _ovlp_1d(pa[x], jx - 2, a, b, A[x] - B[x])
                                  + b * (2 * jx + 1) * S[x]
                                  - 2.0 * b * b * _ovlp_1d(pa[x], jx + 2, a, b, A[x] - B[x]))
                            v += tx * S[(x + 1) % 3] * S[(x + 2) % 3]
                    else:
                        v = 0.0
                        L = li + lj
                        for Z, C in atoms:
                            R = hermite_R(L, p, P - C)
                            acc = 0.0
                            for t in range(pa[0] + pb[0] + 1):
                                for u in range(pa[1] + pb[1] + 1):
                                    for w in range(pa[2] + pb[2] + 1):
                                        acc += E[0][pa[0], pb[0], t] * E[1][pa[1], pb[1], u] * \
                                            E[2][pa[2], pb[2], w] * R[t, u, w]
                            v -= Z * 2.0 * math.pi / p * acc
                    out[ia, jb] += c1 * c2 * v
    return out
