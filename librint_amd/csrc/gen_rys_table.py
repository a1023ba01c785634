"""Generate the piecewise-Chebyshev Rys quadrature table used by the HIP kernels.

Rys quadrature of order n integrates f(t^2) exp(-x t^2) dt over t in [0, 1]
exactly for polynomials f of degree < 2n:

    int_0^1 f(t^2) e^{-x t^2} dt = sum_r w_r f(v_r),    v_r = t_r^2.

The moments of that measure in v are the Boys functions F_k(x), k < 2n.  For
every x the nodes v_r and weights w_r are obtained in 50-digit arithmetic:
moments -> three-term recurrence (Chebyshev algorithm) -> Jacobi matrix ->
symmetric eigenproblem (Golub-Welsch).  The same convention is used by
libcint's CINTrys_roots (reference src/rys_roots.rs:3036-3064), which returns
u = v/(1-v) and the plain weights w with sum_r w_r = F_0(x).

Range [0, XCUT(n)) is split into unit-length segments; on each one every v_r and
w_r is a degree-DEG polynomial in the local variable s in [-1, 1], obtained as
the Chebyshev interpolant at DEG+1 Chebyshev nodes and stored in the monomial
basis of s (converted in 50-digit arithmetic) so that the device can evaluate
it with Estrin's scheme.  Above
XCUT(n) the asymptotic Gauss-Hermite limit is exact to double precision:
v_r = h_r/x, w_r = H_r/sqrt(x), with h_r the squared positive nodes and H_r the
matching weights of the 2n-point Gauss-Hermite rule (weight e^{-s^2}).

Table layout (one block per n, GPU friendly: one lane evaluates all 2n
polynomials of its segment with contiguous loads):
    RYS_CHEB[off(n) + (seg*(DEG+1) + c)*2n + j],  j < n: v_j, j >= n: w_{j-n}
with c the power of s.

Run:  python librint_amd/csrc/gen_rys_table.py         (writes rys_table.inc)
      python librint_amd/csrc/gen_rys_table.py --f32   (writes rys_table_f32.inc:
      degree 5, single precision, for the FP32 recurrence path; same segments
      and cutoffs, arrays named RYS32_*)
The output is committed; the generator needs mpmath and is not run at build time.
"""
import os
import sys
import multiprocessing as mpr

import mpmath as mp

F32 = "--f32" in sys.argv
# roots up to the adjoint of an all-g quartet (l <= 4 raised by 2 on one
# pair: 10 roots); the FP32 table serves the l <= 2 recurrences only
NMAX = 9 if F32 else 10
DEG = 5 if F32 else 11  # polynomial degree per segment
SEGLEN = 1.0      # segment length in x
DPS = 50


def boys(m, x):
    if x == 0:
        return mp.mpf(1) / (2 * m + 1)
    return mp.hyp1f1(m + mp.mpf(1) / 2, m + mp.mpf(3) / 2, -x) / (2 * m + 1)


def rys_exact(n, x):
    """Nodes v (ascending) and weights w of the n-point Rys rule at x."""
    mom = [boys(k, x) for k in range(2 * n)]
    alpha = [mp.mpf(0)] * n
    beta = [mp.mpf(0)] * n
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = list(mom)
    alpha[0] = mom[1] / mom[0]
    beta[0] = mom[0]
    for k in range(1, n):
        sig_new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            sig_new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha[k] = sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1]
        beta[k] = sig_new[k] / sig[k - 1]
        sig_prev, sig = sig, sig_new
    J = mp.zeros(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    idx = sorted(range(n), key=lambda i: E[i])
    return [E[i] for i in idx], [beta[0] * Q[0, i] ** 2 for i in idx]


def hermite_half(n):
    """Squared positive nodes and weights of the 2n-point Gauss-Hermite rule."""
    nodes, weights = [], []
    # roots of H_{2n} by polishing mpmath's polynomial roots
    coeffs = mp.taylor(lambda s: mp.hermite(2 * n, s), 0, 2 * n)[::-1]
    roots = sorted(r.real for r in mp.polyroots(coeffs, maxsteps=500, extraprec=200) if r.real > 0)
    for r in roots:
        r = mp.findroot(lambda s: mp.hermite(2 * n, s), r)
        # w = 2^{N-1} N! sqrt(pi) / (N^2 H_{N-1}(r)^2), N = 2n
        N = 2 * n
        w = mp.mpf(2) ** (N - 1) * mp.factorial(N) * mp.sqrt(mp.pi) / (N * N * mp.hermite(N - 1, r) ** 2)
        nodes.append(r * r)
        weights.append(w)
    return nodes, weights


def asymptotic(n, x, hh):
    h, H = hh
    return [hr / x for hr in h], [Hr / mp.sqrt(x) for Hr in H]


def cheb_fit(job):
    n, seg = job
    mp.mp.dps = DPS
    c = mp.mpf(seg) * SEGLEN + mp.mpf(SEGLEN) / 2
    half = mp.mpf(SEGLEN) / 2
    nodes = [mp.cos(mp.pi * (j + mp.mpf(1) / 2) / (DEG + 1)) for j in range(DEG + 1)]
    vals = [rys_exact(n, c + half * t) for t in nodes]
    # Chebyshev polynomials T_0..T_DEG in the monomial basis (exact integers)
    T = [[mp.mpf(1)] + [mp.mpf(0)] * DEG, [mp.mpf(0), mp.mpf(1)] + [mp.mpf(0)] * (DEG - 1)]
    for m in range(2, DEG + 1):
        T.append([2 * (T[m - 1][k - 1] if k > 0 else 0) - T[m - 2][k] for k in range(DEG + 1)])
    cols = []
    for which in (0, 1):
        for r in range(n):
            f = [vals[j][which][r] for j in range(DEG + 1)]
            cheb = []
            for c_idx in range(DEG + 1):
                a = 2 * mp.fsum(f[j] * mp.cos(mp.pi * c_idx * (j + mp.mpf(1) / 2) / (DEG + 1))
                                for j in range(DEG + 1)) / (DEG + 1)
                cheb.append(a / 2 if c_idx == 0 else a)
            mono = [mp.fsum(cheb[m] * T[m][k] for m in range(DEG + 1)) for k in range(DEG + 1)]
            cols.append([float(x) for x in mono])
    out = [[cols[j][k] for j in range(2 * n)] for k in range(DEG + 1)]
    # validation at off-node points, evaluated in double as the device does (Estrin)
    worst = [0.0, 0.0]
    for t in [-1.0, -0.77, -0.31, 0.13, 0.58, 0.91, 1.0]:
        ex = rys_exact(n, c + half * mp.mpf(t))
        for which in (0, 1):
            for r in range(n):
                j = which * n + r
                v = estrin([out[k][j] for k in range(DEG + 1)], t)
                worst[which] = max(worst[which], abs(v - float(ex[which][r])))
    return n, seg, out, worst


def _fma(a, b, c):
    # exact-rounded a*b + c in double via mpmath (the device uses fma)
    return float(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))


def estrin(a, s):
    """Estrin evaluation of sum_k a[k] s^k (degree 11 or 5), the device's order."""
    if len(a) == 6:
        # single precision: round every step to float32 as the device does
        import numpy as np
        f = np.float32
        a = [f(x) for x in a]
        s = f(s)
        s2 = f(s * s)
        s4 = f(s2 * s2)
        q0, q1, q2 = f(a[1] * s + a[0]), f(a[3] * s + a[2]), f(a[5] * s + a[4])
        return float(f(f(q2 * s4) + f(q1 * s2 + q0)))
    s2 = s * s
    s4 = s2 * s2
    s8 = s4 * s4
    q = [_fma(a[2 * i + 1], s, a[2 * i]) for i in range(6)]
    r0 = _fma(q[1], s2, q[0])
    r1 = _fma(q[3], s2, q[2])
    r2 = _fma(q[5], s2, q[4])
    t0 = _fma(r1, s4, r0)
    return _fma(r2, s8, t0)


def pick_xcut(n, hh):
    """Smallest integer x from which the Hermite limit matches to 1e-16 (v) / 5e-17 (w)."""
    mp.mp.dps = DPS
    x = 20
    while True:
        ok = True
        for xx in (x, x + 0.5, x + 1, x + 3, x + 10):
            v, w = rys_exact(n, mp.mpf(xx))
            va, wa = asymptotic(n, mp.mpf(xx), hh)
            if max(abs(a - b) for a, b in zip(v, va)) > 1e-16 or max(abs(a - b) for a, b in zip(w, wa)) > 5e-17:
                ok = False
                break
        if ok:
            return x
        x += 1


def main():
    mp.mp.dps = DPS
    herm = {n: hermite_half(n) for n in range(1, NMAX + 1)}
    xcut = {n: pick_xcut(n, herm[n]) for n in range(1, NMAX + 1)}
    nseg = {n: int(xcut[n] / SEGLEN) for n in range(1, NMAX + 1)}
    jobs = [(n, s) for n in range(1, NMAX + 1) for s in range(nseg[n])]
    with mpr.Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(cheb_fit, jobs, chunksize=4)
    res.sort()
    off = {}
    flat = []
    worst = {n: [0.0, 0.0] for n in range(1, NMAX + 1)}
    for n in range(1, NMAX + 1):
        off[n] = len(flat)
        for (nn, s, rows, w) in res:
            if nn != n:
                continue
            for row in rows:
                flat.extend(row)
            worst[n][0] = max(worst[n][0], w[0])
            worst[n][1] = max(worst[n][1], w[1])
    if F32:
        write_f32(xcut, nseg, off, herm, flat, worst)
    else:
        write_f64(xcut, nseg, off, herm, flat, worst)
    for n in range(1, NMAX + 1):
        print("n=%d xcut=%d nseg=%d  max|dv|=%.2e max|dw|=%.2e" % (n, xcut[n], nseg[n], worst[n][0], worst[n][1]))


def write_f64(xcut, nseg, off, herm, flat, worst):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rys_table.inc")
    with open(path, "w") as f:
        f.write("// Generated by gen_rys_table.py -- piecewise polynomial Rys nodes/weights\n")
        f.write("// (Chebyshev interpolants stored in the monomial basis of the local variable).\n")
        f.write("// Max abs error of double Estrin evaluation vs %d-digit Golub-Welsch, per n (v, w):\n" % DPS)
        for n in range(1, NMAX + 1):
            f.write("//   n=%d xcut=%d nseg=%d  %.2e  %.2e\n" % (n, xcut[n], nseg[n], worst[n][0], worst[n][1]))
        f.write("// Include after defining RYS_STORAGE (e.g. `static const` on the host,\n")
        f.write("// `__device__ static const` in device code).\n")
        f.write("#define RYS_NMAX %d\n#define RYS_DEG %d\n#define RYS_SEGLEN %.17g\n" % (NMAX, DEG, SEGLEN))
        f.write("RYS_STORAGE int RYS_XCUT[%d] = {0, %s};\n" % (NMAX + 1, ", ".join(str(xcut[n]) for n in range(1, NMAX + 1))))
        f.write("#ifndef RYS_NSEG_DEFINED\n#define RYS_NSEG_DEFINED\n")
        f.write("static constexpr int RYS_NSEG[%d] = {0, %s};\n#endif\n" % (NMAX + 1, ", ".join(str(nseg[n]) for n in range(1, NMAX + 1))))
        f.write("RYS_STORAGE int RYS_OFF[%d] = {0, %s};\n" % (NMAX + 1, ", ".join(str(off[n]) for n in range(1, NMAX + 1))))
        hoff = [0]
        for n in range(1, NMAX + 1):
            hoff.append(hoff[-1] + n)
        f.write("RYS_STORAGE int RYS_HOFF[%d] = {%s};\n" % (NMAX + 1, ", ".join(str(o - 0) for o in [0] + hoff[:-1])))
        f.write("RYS_STORAGE double RYS_HERM_NODE2[%d] = {\n" % hoff[-1])
        for n in range(1, NMAX + 1):
            f.write("  " + ", ".join("%.17g" % float(v) for v in herm[n][0]) + ",\n")
        f.write("};\nRYS_STORAGE double RYS_HERM_W[%d] = {\n" % hoff[-1])
        for n in range(1, NMAX + 1):
            f.write("  " + ", ".join("%.17g" % float(v) for v in herm[n][1]) + ",\n")
        f.write("};\n#define RYS_CHEB_LEN %d\n" % len(flat))
        f.write("RYS_STORAGE double RYS_CHEB[RYS_CHEB_LEN] = {\n")
        for i in range(0, len(flat), 4):
            f.write("  " + ", ".join("%.17g" % v for v in flat[i:i + 4]) + ",\n")
        f.write("};\n")


def write_f32(xcut, nseg, off, herm, flat, worst):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rys_table_f32.inc")
    with open(path, "w") as f:
        f.write("// Generated by gen_rys_table.py --f32 -- single-precision piecewise polynomial\n")
        f.write("// Rys nodes/weights (degree %d in the local variable, monomial basis) for the\n" % DEG)
        f.write("// FP32 recurrence path.  Max abs error of float Estrin evaluation vs %d-digit\n" % DPS)
        f.write("// Golub-Welsch, per n (v, w):\n")
        for n in range(1, NMAX + 1):
            f.write("//   n=%d  %.2e  %.2e\n" % (n, worst[n][0], worst[n][1]))
        f.write("// Same segments and large-x cutoffs as rys_table.inc; include after defining\n")
        f.write("// RYS_STORAGE.\n#define RYS32_DEG %d\n" % DEG)
        f.write("RYS_STORAGE int RYS32_OFF[%d] = {0, %s};\n" % (NMAX + 1, ", ".join(str(off[n]) for n in range(1, NMAX + 1))))
        f.write("#define RYS32_LEN %d\n" % len(flat))
        f.write("RYS_STORAGE float RYS32_TAB[RYS32_LEN] = {\n")
        for i in range(0, len(flat), 6):
            f.write("  " + ", ".join("%.9ef" % v for v in flat[i:i + 6]) + ",\n")
        f.write("};\n")


if __name__ == "__main__":
    sys.exit(main())
