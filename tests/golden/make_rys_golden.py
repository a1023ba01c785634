"""Write tests/golden/rys_mpmath.json: Boys values and Rys nodes/weights at
50 digits (independent of both the oracle and the GPU table generator's code
path: the rule here comes from mpmath's own Gauss quadrature of the moment
problem via a Lanczos process on a high-order Gauss-Legendre discretisation).

Run: python tests/golden/make_rys_golden.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 50
XS = [0.0, 1e-9, 3e-7, 1e-4, 0.05, 0.5, 1.0, 2.3, 4.9, 7.5, 11.0, 15.5, 19.99, 24.2, 29.7,
      34.9, 35.1, 42.5, 43.5, 49.9, 50.5, 55.5, 56.5, 60.7, 61.3, 66.6, 72.1, 80.0, 90.0, 150.0, 1e3, 1e5]


def boys(m, x):
    if x == 0:
        return mp.mpf(1) / (2 * m + 1)
    return mp.gammainc(m + mp.mpf(1) / 2, 0, x) / (2 * x ** (m + mp.mpf(1) / 2))


def rys(n, x, N=240):
    """Lanczos on the discrete measure of an N-point Gauss-Legendre rule in t."""
    T = mp.mpf(1) if x <= 100 else mp.sqrt(mp.mpf(200) / x)
    nodes, weights = [], []
    for t, w in zip(*mp_gauss_legendre(N)):
        tt = T * (t + 1) / 2
        nodes.append(tt * tt)
        weights.append(T / 2 * w * mp.e ** (-x * tt * tt))
    # Stieltjes procedure
    p_prev = [mp.mpf(0)] * N
    p = [mp.mpf(1)] * N
    nrm = mp.fsum(weights)
    b0 = nrm
    alpha, beta = [], [b0]
    for k in range(n):
        a = mp.fsum(w * v * q * q for w, v, q in zip(weights, nodes, p)) / nrm
        alpha.append(a)
        if k + 1 == n:
            break
        pn = [(v - a) * q - (beta[k] if k > 0 else 0) * qp for v, q, qp in zip(nodes, p, p_prev)]
        nn = mp.fsum(w * q * q for w, q in zip(weights, pn))
        beta.append(nn / nrm)
        p_prev, p, nrm = p, pn, nn
    J = mp.zeros(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    idx = sorted(range(n), key=lambda i: E[i])
    return [E[i] for i in idx], [b0 * Q[0, i] ** 2 for i in idx]


_GL = {}


def mp_gauss_legendre(N):
    if N not in _GL:
        xs, ws = [], []
        for i in range(1, N + 1):
            z = mp.cos(mp.pi * (i - mp.mpf(1) / 4) / (N + mp.mpf(1) / 2))
            for _ in range(100):
                p0, p1 = mp.mpf(1), z
                for k in range(2, N + 1):
                    p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
                dp = N * (z * p1 - p0) / (z * z - 1)
                dz = p1 / dp
                z -= dz
                if abs(dz) < mp.mpf(10) ** (-45):
                    break
            p0, p1 = mp.mpf(1), z
            for k in range(2, N + 1):
                p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
            dp = N * (z * p1 - p0) / (z * z - 1)
            xs.append(z)
            ws.append(2 / ((1 - z * z) * dp * dp))
        _GL[N] = (xs, ws)
    return _GL[N]


def main():
    out = {"boys": [], "rys": []}
    for x in XS:
        out["boys"].append({"x": x, "F": [float(boys(m, mp.mpf(x))) for m in range(0, 19)]})
    for n in range(1, 11):
        for x in XS:
            if x > 2000:
                continue
            v, w = rys(n, mp.mpf(x))
            out["rys"].append({"n": n, "x": x, "v": [float(a) for a in v], "w": [float(a) for a in w]})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rys_mpmath.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", path, len(out["rys"]))


if __name__ == "__main__":
    main()
