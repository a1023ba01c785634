// One-electron input staging and the host-side derivatives of the
// one-electron integrals (basis parameters and nuclear coordinates) with the
// semantics of the reference's cint1e_{ovlp,kin,nuc}_cart
// (src/cint1e.rs:925-1025, src/intor1.rs:168-218, src/g1e.rs:241-395):
// coefficients as stored in env times common_fac_sp(l), point nuclei,
// V = sum_C -Z_C <i|1/|r-C||j>, T = <i|-1/2 nabla^2|j>.  The matrices
// themselves are computed on the GPU (int1e_kernel.hip).  S and T derivatives use
// the Obara-Saika 1D overlap recursion, V derivatives Rys quadrature.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "int1e.h"
#include "rys_eval.h"

namespace librint {

static double fac_sp(int l)
{
    if (l == 0) return 0.282094791773878143;
    if (l == 1) return 0.488602511902919921;
    return 1.0;
}

static void cart_pows(int l, int *nx, int *ny, int *nz)
{
    int n = 0;
    for (int lx = l; lx >= 0; lx--)
        for (int ly = l - lx; ly >= 0; ly--) {
            nx[n] = lx;
            ny[n] = ly;
            nz[n] = l - lx - ly;
            n++;
        }
}

// s[i][j], i <= imax, j <= jmax  (1D overlap of x_A^i x_B^j times the Gaussian)
static void os_1d(int imax, int jmax, double PA, double PB, double p, double s00, double s[8][10])
{
    const double h = 0.5 / p;
    for (int i = 0; i <= imax; i++)
        for (int j = 0; j <= jmax; j++) {
            double v;
            if (i == 0 && j == 0) v = s00;
            else if (i > 0) {
                v = PA * s[i - 1][j];
                if (i > 1) v += (i - 1) * h * s[i - 2][j];
                if (j > 0) v += j * h * s[i - 1][j - 1];
            } else {
                v = PB * s[i][j - 1];
                if (j > 1) v += (j - 1) * h * s[i][j - 2];
            }
            s[i][j] = v;
        }
}

Int1eInput make_int1e_input(const int *atm, size_t atm_l, const int *bas, size_t bas_l, const double *env)
{
    Int1eInput in;
    in.env.assign(env, env + 0);
    size_t env_need = 0;
    const int nb = (int)(bas_l / 8);
    int ao = 0;
    for (int s = 0; s < nb; s++) {
        const int *b = bas + 8 * s;
        for (int c = 0; c < b[3]; c++) {
            Shell1 sh;
            sh.l = b[1];
            sh.nprim = b[2];
            sh.pexp = b[5];
            sh.pcoef = b[6] + c * b[2];
            sh.atom = b[0];
            const int pr = atm[6 * b[0] + 1];
            for (int d = 0; d < 3; d++) sh.r[d] = env[pr + d];
            sh.ao = ao;
            ao += (sh.l + 1) * (sh.l + 2) / 2;
            env_need = std::max(env_need, (size_t)std::max(sh.pexp, sh.pcoef) + sh.nprim);
            in.shells.push_back(sh);
        }
    }
    for (size_t a = 0; a < atm_l / 6; a++) {
        const int pr = atm[6 * a + 1];
        env_need = std::max(env_need, (size_t)pr + 3);
        in.charges.push_back(std::fabs((double)atm[6 * a]));
        in.centers.push_back({env[pr], env[pr + 1], env[pr + 2]});
    }
    in.env.assign(env, env + env_need);
    in.nao = ao;
    return in;
}

// ---------------------------------------------------------------------------
// Parameter and nuclear derivatives (adjoint of the 1e part of the energy).
//
// For a primitive x_A^a exp(-alpha x_A^2) on the bra:
//   d/dA_x   = 2 alpha G(a + 1_x) - a_x G(a - 1_x)
//   d/dalpha = -(G(a + 2_x) + G(a + 2_y) + G(a + 2_z))
//   d/dc     = primitive integral / c
// With W and O symmetric, the ket-side derivative of O_ab equals the bra-side
// derivative of O_ba, so sum_ab W_ab dO_ab = 2 sum_ab W_ab d_bra O_ab.  For V
// the point-charge position C enters through A - C and B - C only, so
// d/dC V^C_ab = -(d/dA + d/dB) V^C_ab.

namespace {

struct PrimTables {
    // 1D tables for bra powers up to li + 2 and ket powers up to lj
    double s[3][8][10];
    double t[3][8][8];
};

// value of a cartesian S (typ 0) or T (typ 1) primitive integral from the 1D tables
inline double st_value(const PrimTables &T, int typ, const int *ia, const int *jb)
{
    const double sx = T.s[0][ia[0]][jb[0]], sy = T.s[1][ia[1]][jb[1]], sz = T.s[2][ia[2]][jb[2]];
    if (typ == 0) return sx * sy * sz;
    return T.t[0][ia[0]][jb[0]] * sy * sz + sx * T.t[1][ia[1]][jb[1]] * sz + sx * sy * T.t[2][ia[2]][jb[2]];
}

// bra derivatives of one primitive integral: value, exponent, A-gradient
template <typename F>
inline void bra_derivs(F val, const int *ia, const int *jb, double alpha, double &v0, double &de, double dA[3])
{
    v0 = val(ia, jb);
    de = 0.0;
    for (int d = 0; d < 3; d++) {
        int up[3] = {ia[0], ia[1], ia[2]};
        up[d] += 2;
        de -= val(up, jb);
        int p1[3] = {ia[0], ia[1], ia[2]};
        p1[d] += 1;
        double g = 2.0 * alpha * val(p1, jb);
        if (ia[d] > 0) {
            int m1[3] = {ia[0], ia[1], ia[2]};
            m1[d] -= 1;
            g -= ia[d] * val(m1, jb);
        }
        dA[d] = g;
    }
}

}  // namespace

// Visit every ordered shell pair and primitive pair; f(A, ip, B, jp, fa, fb, v0, de, dA, C)
// receives the bra derivatives of the primitive integral (C = -1 for S/T, the
// charge index for the V contribution of that charge).  Coefficients are
// folded in except the bra coefficient (v0 is d/dc_bra).
template <typename Visit>
static void visit_bra_derivs(const Int1eInput &in, int typ, Visit f)
{
    const int ns = (int)in.shells.size();
    // the 1D tables below hold the bra momentum raised by 2 for l <= 2
    for (const Shell1 &S : in.shells)
        if (S.l < 0 || S.l > 2) throw std::invalid_argument("one-electron derivatives cover l <= 2");
    for (int ish = 0; ish < ns; ish++)
        for (int jsh = 0; jsh < ns; jsh++) {
            const Shell1 &A = in.shells[ish], &B = in.shells[jsh];
            const int li = A.l, lj = B.l;
            if (li > 3 || lj > 3) throw std::invalid_argument("l > 3 not supported for one-electron integrals");
            const int nfi = (li + 1) * (li + 2) / 2, nfj = (lj + 1) * (lj + 2) / 2;
            int inx[15], iny[15], inz[15], jnx[15], jny[15], jnz[15];
            cart_pows(li, inx, iny, inz);
            cart_pows(lj, jnx, jny, jnz);
            double AB[3], AB2 = 0;
            for (int d = 0; d < 3; d++) {
                AB[d] = A.r[d] - B.r[d];
                AB2 += AB[d] * AB[d];
            }
            const double cf = fac_sp(li) * fac_sp(lj);
            for (int jp = 0; jp < B.nprim; jp++)
                for (int ip = 0; ip < A.nprim; ip++) {
                    const double a = in.env[A.pexp + ip], b = in.env[B.pexp + jp];
                    const double cb = in.env[B.pcoef + jp];
                    const double p = a + b;
                    double P[3];
                    for (int d = 0; d < 3; d++) P[d] = (a * A.r[d] + b * B.r[d]) / p;
                    const double pref = cb * cf * std::exp(-a * b / p * AB2);  // bra coefficient left out
                    if (typ <= 1) {
                        PrimTables T;
                        for (int d = 0; d < 3; d++) {
                            os_1d(li + 2, lj + 2, P[d] - A.r[d], P[d] - B.r[d], p, std::sqrt(M_PI / p), T.s[d]);
                            for (int i = 0; i <= li + 2; i++)
                                for (int j = 0; j <= lj; j++) {
                                    double tt = -2.0 * b * (2 * j + 1) * T.s[d][i][j] + 4.0 * b * b * T.s[d][i][j + 2];
                                    if (j >= 2) tt += j * (j - 1) * T.s[d][i][j - 2];
                                    T.t[d][i][j] = -0.5 * tt;
                                }
                        }
                        auto val = [&](const int *ia, const int *jb) { return pref * st_value(T, typ, ia, jb); };
                        for (int jf = 0; jf < nfj; jf++)
                            for (int fi = 0; fi < nfi; fi++) {
                                const int ia[3] = {inx[fi], iny[fi], inz[fi]}, jb[3] = {jnx[jf], jny[jf], jnz[jf]};
                                double v0, de, dA[3];
                                bra_derivs(val, ia, jb, a, v0, de, dA);
                                f(ish, ip, jsh, fi, jf, v0, de, dA, -1);
                            }
                    } else {
                        const int ltot = li + 2 + lj;
                        const int nr = ltot / 2 + 1;
                        for (size_t c = 0; c < in.charges.size(); c++) {
                            const double *C = in.centers[c].data();
                            double PC2 = 0;
                            for (int d = 0; d < 3; d++) PC2 += (P[d] - C[d]) * (P[d] - C[d]);
                            double v[9], w[9];
                            rys_nodes_n(nr, p * PC2, v, w);
                            const double fac = -in.charges[c] * 2.0 * M_PI / p * pref;
                            // I[r][d][i][j], bra powers up to li + 2
                            double I[5][3][8][4];
                            for (int r = 0; r < nr; r++)
                                for (int d = 0; d < 3; d++) {
                                    const double c00 = (P[d] - A.r[d]) - v[r] * (P[d] - C[d]);
                                    const double b10 = (1 - v[r]) * 0.5 / p;
                                    double g[12];
                                    g[0] = 1;
                                    if (ltot > 0) g[1] = c00;
                                    for (int i = 1; i < ltot; i++) g[i + 1] = c00 * g[i] + i * b10 * g[i - 1];
                                    double H[12][4];
                                    for (int i = 0; i <= ltot; i++) H[i][0] = g[i];
                                    for (int j = 1; j <= lj; j++)
                                        for (int i = 0; i <= ltot - j; i++) H[i][j] = H[i + 1][j - 1] + AB[d] * H[i][j - 1];
                                    for (int i = 0; i <= li + 2; i++)
                                        for (int j = 0; j <= lj; j++) I[r][d][i][j] = H[i][j];
                                }
                            auto val = [&](const int *ia, const int *jb) {
                                double acc = 0;
                                for (int r = 0; r < nr; r++)
                                    acc += w[r] * I[r][0][ia[0]][jb[0]] * I[r][1][ia[1]][jb[1]] * I[r][2][ia[2]][jb[2]];
                                return fac * acc;
                            };
                            for (int jf = 0; jf < nfj; jf++)
                                for (int fi = 0; fi < nfi; fi++) {
                                    const int ia[3] = {inx[fi], iny[fi], inz[fi]}, jb[3] = {jnx[jf], jny[jf], jnz[jf]};
                                    double v0, de, dA[3];
                                    bra_derivs(val, ia, jb, a, v0, de, dA);
                                    f(ish, ip, jsh, fi, jf, v0, de, dA, (int)c);
                                }
                        }
                    }
                }
        }
}

void int1e_grad(const Int1eInput &in, int typ, const double *W, double *gatom, double *denv, size_t env_l)
{
    const int n = in.nao;
    visit_bra_derivs(in, typ, [&](int ish, int ip, int jsh, int fi, int jf, double v0, double de, const double *dA, int c) {
        const Shell1 &A = in.shells[ish], &B = in.shells[jsh];
        const double w2 = 2.0 * W[(size_t)(A.ao + fi) * n + B.ao + jf];
        if (w2 == 0.0) return;
        const double ca = in.env[A.pcoef + ip];
        if (denv) {
            if ((size_t)(A.pexp + ip) < env_l) denv[A.pexp + ip] += w2 * ca * de;
            if ((size_t)(A.pcoef + ip) < env_l) denv[A.pcoef + ip] += w2 * v0;
        }
        if (gatom)
            for (int d = 0; d < 3; d++) {
                const double g = w2 * ca * dA[d];
                gatom[3 * A.atom + d] += g;
                if (c >= 0) gatom[3 * c + d] -= g;
            }
    });
}

void overlap_param_derivs(const Int1eInput &in, size_t s1, size_t L, double *dS)
{
    const int n = in.nao;
    std::memset(dS, 0, sizeof(double) * (size_t)n * n * L);
    visit_bra_derivs(in, 0, [&](int ish, int ip, int jsh, int fi, int jf, double v0, double de, const double *, int) {
        const Shell1 &A = in.shells[ish], &B = in.shells[jsh];
        const size_t a = A.ao + fi, b = B.ao + jf;
        const double ca = in.env[A.pcoef + ip];
        const size_t ie = (size_t)(A.pexp + ip), ic = (size_t)(A.pcoef + ip);
        // d_bra S_ab is also the ket derivative of S_ba
        if (ie >= s1 && ie < s1 + L) {
            dS[(a * n + b) * L + (ie - s1)] += ca * de;
            dS[(b * n + a) * L + (ie - s1)] += ca * de;
        }
        if (ic >= s1 && ic < s1 + L) {
            dS[(a * n + b) * L + (ic - s1)] += v0;
            dS[(b * n + a) * L + (ic - s1)] += v0;
        }
    });
}

void nuclear_repulsion_grad(const Int1eInput &in, double *gatom)
{
    const size_t na = in.charges.size();
    for (size_t i = 0; i < na; i++)
        for (size_t j = 0; j < na; j++) {
            if (i == j) continue;
            double d[3], r2 = 0;
            for (int k = 0; k < 3; k++) {
                d[k] = in.centers[i][k] - in.centers[j][k];
                r2 += d[k] * d[k];
            }
            const double f = in.charges[i] * in.charges[j] / (r2 * std::sqrt(r2));
            for (int k = 0; k < 3; k++) gatom[3 * i + k] -= f * d[k];
        }
}

double nuclear_repulsion(const Int1eInput &in)
{ // reference src/scf.rs:946-953 (point charges)
    double E = 0;
    for (size_t i = 0; i < in.charges.size(); i++)
        for (size_t j = 0; j < i; j++) {
            double r2 = 0;
            for (int d = 0; d < 3; d++) r2 += (in.centers[i][d] - in.centers[j][d]) * (in.centers[i][d] - in.centers[j][d]);
            E += in.charges[i] * in.charges[j] / std::sqrt(r2);
        }
    return E;
}

}  // namespace librint
