// RHF driver on top of the GPU Fock build.
//
// Reference mode (default, reference src/scf.rs:743-799): P0 = identity,
// plain Roothaan iteration F -> X^T F X -> eigh -> aufbau P, stop when
// |P - P_old|_F / 2 <= conv, then the validity checks of src/scf.rs:572-600
// and the energy E = tr(P(T+V)) + 1/2 tr(P G[P]) + E_nuc (src/scf.rs:827-956).
// Options: core-Hamiltonian or caller-supplied starting density, and Pulay
// DIIS extrapolation (not in the reference; off unless requested).
#include "scf.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "int1e_gpu.h"
#include "linalg.h"

namespace librint {

static bool solve_linear(int n, std::vector<double> A, std::vector<double> b, std::vector<double> &x)
{
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(A[r * n + c]) > std::fabs(A[p * n + c])) p = r;
        if (std::fabs(A[p * n + c]) < 1e-300) return false;
        if (p != c) {
            for (int k = 0; k < n; k++) std::swap(A[c * n + k], A[p * n + k]);
            std::swap(b[c], b[p]);
        }
        for (int r = c + 1; r < n; r++) {
            const double f = A[r * n + c] / A[c * n + c];
            for (int k = c; k < n; k++) A[r * n + k] -= f * A[c * n + k];
            b[r] -= f * b[c];
        }
    }
    x.assign(n, 0.0);
    for (int r = n - 1; r >= 0; r--) {
        double s = b[r];
        for (int k = r + 1; k < n; k++) s -= A[r * n + k] * x[k];
        x[r] = s / A[r * n + r];
    }
    return true;
}

double energy_fast(Context &ctx, const Int1eInput &in, const double *P, const double *H)
{
    const int n = ctx.nao();
    const size_t nn = (size_t)n * n;
    std::vector<double> G(nn);
    ctx.fock_host(P, G.data());
    double E = 0;
    for (size_t i = 0; i < nn; i++) E += P[i] * (H[i] + 0.5 * G[i]);
    return E + nuclear_repulsion(in);
}

ScfOut scf_run(Context &ctx, const Int1eInput &in, int nelec, int imax, double conv, const ScfOptions &o)
{
    ScfOut out;
    const int n = ctx.nao();
    const size_t nn = (size_t)n * n;
    if (in.nao != n) throw std::runtime_error("basis mismatch between 1e and 2e inputs");
    if (nelec % 2) throw std::invalid_argument("closed-shell RHF needs an even electron count");
    if (nelec / 2 > n) throw std::invalid_argument("more occupied orbitals than basis functions");
    std::vector<double> S(nn), H(nn), V(nn);
    int1e_gpu(in, S.data(), H.data(), V.data());
    for (size_t i = 0; i < nn; i++) H[i] += V[i];

    std::vector<double> s(n), U(nn), X(nn), tmp(nn), Fp(nn), C(nn), F(nn), G(nn), P(nn), Pold(nn), eps(n);
    eigh(n, S.data(), s.data(), U.data());
    if (s[0] < 1e-8) {
        out.status = 3;
        out.msg = "overlap matrix is near-singular (smallest eigenvalue < 1e-8)";
        return out;
    }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) X[(size_t)i * n + k] = U[(size_t)i * n + k] / std::sqrt(s[k]);
    const int nocc = nelec / 2;
    auto density_from_F = [&](const std::vector<double> &Fm) {
        gemm_tn(n, X.data(), Fm.data(), tmp.data());  // X^T F
        gemm_nn(n, tmp.data(), X.data(), Fp.data());  // X^T F X
        eigh(n, Fp.data(), eps.data(), tmp.data());
        gemm_nn(n, X.data(), tmp.data(), C.data());
        for (int mu = 0; mu < n; mu++)
            for (int nu = 0; nu < n; nu++) {
                double acc = 0;
                for (int i = 0; i < nocc; i++) acc += C[(size_t)mu * n + i] * C[(size_t)nu * n + i];
                P[(size_t)mu * n + nu] = 2.0 * acc;
            }
    };
    if (o.guess == 0) {
        std::fill(P.begin(), P.end(), 0.0);
        for (int i = 0; i < n; i++) P[(size_t)i * n + i] = 1.0;
    } else if (o.guess == 1) {
        density_from_F(H);
    } else {
        if (!o.P0) throw std::invalid_argument("guess=2 needs a starting density");
        std::copy(o.P0, o.P0 + nn, P.begin());
    }
    std::vector<std::vector<double>> dF, dE;
    double delta = 1.0, Eprev = 0.0, dEn = 1.0;
    const double Enuc = o.conv_e > 0.0 ? nuclear_repulsion(in) : 0.0;
    int it = 0;
    auto converged = [&]() { return delta <= conv && (o.conv_e <= 0.0 || std::fabs(dEn) <= o.conv_e); };
    while (!converged() && it < imax) {
        Pold = P;
        ctx.fock_host(P.data(), G.data());
        for (size_t i = 0; i < nn; i++) F[i] = H[i] + G[i];
        if (o.conv_e > 0.0) {
            // energy of the density entering this iteration
            double Ek = Enuc;
            for (size_t i = 0; i < nn; i++) Ek += P[i] * (H[i] + 0.5 * G[i]);
            dEn = it == 0 ? 1.0 : Ek - Eprev;
            Eprev = Ek;
        }
        if (o.diis) {
            // e = X^T (F P S - S P F) X
            std::vector<double> FP(nn), FPS(nn), e(nn), t2(nn);
            gemm_nn(n, F.data(), P.data(), FP.data());
            gemm_nn(n, FP.data(), S.data(), FPS.data());
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++) e[(size_t)i * n + j] = FPS[(size_t)i * n + j] - FPS[(size_t)j * n + i];
            gemm_tn(n, X.data(), e.data(), t2.data());
            gemm_nn(n, t2.data(), X.data(), e.data());
            dF.push_back(F);
            dE.push_back(e);
            if ((int)dF.size() > 8) {
                dF.erase(dF.begin());
                dE.erase(dE.begin());
            }
            const int m = (int)dF.size();
            if (m >= 2) {
                std::vector<double> B((m + 1) * (m + 1), 0.0), rhs(m + 1, 0.0), c;
                for (int i = 0; i < m; i++)
                    for (int j = 0; j <= i; j++) {
                        double d = 0;
                        for (size_t k = 0; k < nn; k++) d += dE[i][k] * dE[j][k];
                        B[i * (m + 1) + j] = B[j * (m + 1) + i] = d;
                    }
                for (int i = 0; i < m; i++) B[i * (m + 1) + m] = B[m * (m + 1) + i] = -1.0;
                rhs[m] = -1.0;
                if (solve_linear(m + 1, B, rhs, c)) {
                    std::fill(F.begin(), F.end(), 0.0);
                    for (int i = 0; i < m; i++)
                        for (size_t k = 0; k < nn; k++) F[k] += c[i] * dF[i][k];
                }
            }
        }
        density_from_F(F);
        delta = 0;
        for (size_t i = 0; i < nn; i++) delta += (P[i] - Pold[i]) * (P[i] - Pold[i]);
        delta = std::sqrt(delta) / 2.0;
        it++;
        if (o.trace) std::fprintf(stderr, "librint scf iter %d delta %.3e\n", it, delta);
    }
    out.iters = it;
    out.delta = delta;
    out.dE = dEn;
    out.P = P;
    if (!converged()) {
        out.status = 1;
        char buf[160];
        std::snprintf(buf, sizeof buf, "SCF did not converge in %d cycles: |dP| = %.3e > %.3e", imax, delta, conv);
        out.msg = buf;
        return out;
    }
    // validity checks (reference src/scf.rs:572-600)
    double tr = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) tr += P[(size_t)i * n + j] * S[(size_t)j * n + i];
    if (std::fabs(tr - nelec) > 1e-6) {
        out.status = 2;
        out.msg = "converged density has tr(PS) != nelec";
        return out;
    }
    gemm_nn(n, P.data(), S.data(), tmp.data());
    gemm_nn(n, tmp.data(), P.data(), Fp.data());
    double idem = 0;
    for (size_t i = 0; i < nn; i++) idem = std::max(idem, std::fabs(Fp[i] - 2.0 * P[i]));
    if (idem > 1e-6) {
        out.status = 2;
        out.msg = "converged density is not idempotent";
        return out;
    }
    out.E = energy_fast(ctx, in, P.data(), H.data());
    out.status = 0;
    return out;
}

}  // namespace librint
