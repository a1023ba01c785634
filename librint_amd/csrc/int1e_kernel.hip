// One-electron integrals S, T, V on the GPU (cartesian, l <= 4): semantics of
// the reference's cint1e_{ovlp,kin,nuc}_cart (src/cint1e.rs:925-1025,
// src/intor1.rs:168-218, src/g1e.rs:241-395): coefficients as stored in env
// times common_fac_sp(l), point nuclei, V = sum_C -Z_C <i|1/|r-C||j>,
// T = <i|-1/2 nabla^2|j>.
//
// One thread per ordered shell pair (i, j) of one angular class (LI, LJ),
// writing the (i, j) block of each requested matrix.  S and T use the
// Obara-Saika 1D overlaps (T from the second derivative on the ket), V the
// Rys quadrature with the same tables as the ERI kernels.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "int1e_gpu.h"
#include "rys_eval.h"

namespace librint {

// cartesian exponents of component f of a shell of angular momentum l, in
// libcint's order (x exponent descending, then y descending)
__host__ __device__ constexpr int c1x(int l, int f)
{
    int n = 0;
    for (int lx = l; lx >= 0; lx--)
        for (int ly = l - lx; ly >= 0; ly--) {
            if (n == f) return lx;
            n++;
        }
    return 0;
}
__host__ __device__ constexpr int c1y(int l, int f)
{
    int n = 0;
    for (int lx = l; lx >= 0; lx--)
        for (int ly = l - lx; ly >= 0; ly--) {
            if (n == f) return ly;
            n++;
        }
    return 0;
}
__host__ __device__ constexpr int c1e(int l, int f, int d)
{
    return d == 0 ? c1x(l, f) : d == 1 ? c1y(l, f) : l - c1x(l, f) - c1y(l, f);
}

__device__ __forceinline__ double fac_sp_dev(int l)
{
    return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0);
}

template <int LI, int LJ>
__global__ __launch_bounds__(128) void int1e_kernel(const int2 *__restrict__ pairs, int npair,
                                                    const Shell1Dev *__restrict__ sh, const double *__restrict__ env,
                                                    const double4 *__restrict__ nuc, int nnuc, int nao, int mask,
                                                    double *__restrict__ S, double *__restrict__ T,
                                                    double *__restrict__ V)
{
    constexpr int NFI = (LI + 1) * (LI + 2) / 2, NFJ = (LJ + 1) * (LJ + 2) / 2;
    constexpr int NR = (LI + LJ) / 2 + 1;
    constexpr int LT = LI + LJ;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npair) return;
    const int2 pr = pairs[t];
    const Shell1Dev A = sh[pr.x], B = sh[pr.y];
    double AB[3], AB2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        AB[d] = A.r[d] - B.r[d];
        AB2 += AB[d] * AB[d];
    }
    const double cf = fac_sp_dev(LI) * fac_sp_dev(LJ);
    double bs[NFI * NFJ], bt[NFI * NFJ], bv[NFI * NFJ];
#pragma unroll
    for (int e = 0; e < NFI * NFJ; e++) bs[e] = bt[e] = bv[e] = 0.0;
    for (int jp = 0; jp < B.nprim; jp++)
        for (int ip = 0; ip < A.nprim; ip++) {
            const double a = env[A.pexp + ip], b = env[B.pexp + jp];
            const double p = a + b, h = 0.5 / p;
            double P[3];
#pragma unroll
            for (int d = 0; d < 3; d++) P[d] = (a * A.r[d] + b * B.r[d]) / p;
            const double pref = env[A.pcoef + ip] * env[B.pcoef + jp] * cf * exp(-a * b / p * AB2);
            if (mask & 3) {
                // s[d][i][j]: 1D overlap of x_A^i x_B^j, i <= LI, j <= LJ + 2
                double s[3][LI + 1][LJ + 3];
                const double s00 = sqrt(M_PI / p);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double PA = P[d] - A.r[d], PB = P[d] - B.r[d];
#pragma unroll
                    for (int i = 0; i <= LI; i++)
#pragma unroll
                        for (int j = 0; j <= LJ + 2; j++) {
                            double v;
                            if (i == 0 && j == 0) v = s00;
                            else if (i > 0) {
                                v = PA * s[d][i - 1][j];
                                if (i > 1) v += (i - 1) * h * s[d][i - 2][j];
                                if (j > 0) v += j * h * s[d][i - 1][j - 1];
                            } else {
                                v = PB * s[d][i][j - 1];
                                if (j > 1) v += (j - 1) * h * s[d][i][j - 2];
                            }
                            s[d][i][j] = v;
                        }
                }
#pragma unroll
                for (int fj = 0; fj < NFJ; fj++)
#pragma unroll
                    for (int fi = 0; fi < NFI; fi++) {
                        double sv[3], tv[3];
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            const int i = c1e(LI, fi, d), j = c1e(LJ, fj, d);
                            sv[d] = s[d][i][j];
                            double tt = -2.0 * b * (2 * j + 1) * s[d][i][j] + 4.0 * b * b * s[d][i][j + 2];
                            if (j >= 2) tt += j * (j - 1) * s[d][i][j - 2];
                            tv[d] = -0.5 * tt;
                        }
                        bs[fi + NFI * fj] += pref * sv[0] * sv[1] * sv[2];
                        bt[fi + NFI * fj] += pref * (tv[0] * sv[1] * sv[2] + sv[0] * tv[1] * sv[2] + sv[0] * sv[1] * tv[2]);
                    }
            }
            if (mask & 4) {
                for (int c = 0; c < nnuc; c++) {
                    const double4 C = nuc[c];
                    const double PC[3] = {P[0] - C.x, P[1] - C.y, P[2] - C.z};
                    const double PC2 = PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2];
                    double rv[NR], rw[NR];
                    rys_nodes<NR>(p * PC2, rv, rw);
                    const double fac = -C.w * 2.0 * M_PI / p * pref;
#pragma unroll
                    for (int r = 0; r < NR; r++) {
                        // I[d][i][j], i <= LI, j <= LJ, from g(i + j) and the HRR
                        double I[3][LT + 1][LJ + 1];
                        const double b10 = (1.0 - rv[r]) * h;
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            const double c00 = (P[d] - A.r[d]) - rv[r] * PC[d];
                            I[d][0][0] = 1.0;
                            if (LT > 0) I[d][1][0] = c00;
#pragma unroll
                            for (int i = 1; i < LT; i++) I[d][i + 1][0] = c00 * I[d][i][0] + i * b10 * I[d][i - 1][0];
#pragma unroll
                            for (int j = 1; j <= LJ; j++)
#pragma unroll
                                for (int i = 0; i <= LT - j; i++) I[d][i][j] = I[d][i + 1][j - 1] + AB[d] * I[d][i][j - 1];
                        }
                        const double fw = fac * rw[r];
#pragma unroll
                        for (int fj = 0; fj < NFJ; fj++)
#pragma unroll
                            for (int fi = 0; fi < NFI; fi++)
                                bv[fi + NFI * fj] += fw * I[0][c1e(LI, fi, 0)][c1e(LJ, fj, 0)] *
                                                     I[1][c1e(LI, fi, 1)][c1e(LJ, fj, 1)] *
                                                     I[2][c1e(LI, fi, 2)][c1e(LJ, fj, 2)];
                    }
                }
            }
        }
    const size_t n = (size_t)nao;
#pragma unroll
    for (int fj = 0; fj < NFJ; fj++)
#pragma unroll
        for (int fi = 0; fi < NFI; fi++) {
            const size_t o = (size_t)(A.ao + fi) * n + B.ao + fj;
            if (S) S[o] = bs[fi + NFI * fj];
            if (T) T[o] = bt[fi + NFI * fj];
            if (V) V[o] = bv[fi + NFI * fj];
        }
}

// ------------------------------------------------------------ derivatives --
// Bra derivatives of the primitive integrals of one ordered shell pair (the
// semantics of the reference's dscf.rs one-electron parts: src/dscf.rs:96-150
// for dS, :320-336 for dHcore).  With W and O symmetric, the ket derivative
// of O_ab is the bra derivative of O_ba, so sum_ab W_ab dO_ab = 2 sum_ab W_ab
// d_bra O_ab; a point charge C enters V^C only through A - C and B - C, so
// d/dC V^C = -(d/dA + d/dB) V^C.  Per primitive x_A^a exp(-alpha r_A^2):
//   d/dA_x = 2 alpha G(a + 1_x) - a_x G(a - 1_x),  d/dalpha = -sum_x G(a + 2_x),
//   d/dc = G / c.
// The 1D tables hold bra powers up to LI + 2.
//
// S and T: one thread per ordered shell pair.  V: a segment of CW = 2^k
// lanes per pair (CW = the point-charge count rounded up to a power of two,
// at most 64); lane cs of a segment takes the charges cs, cs + CW, ...  A
// segment shares its pair's primitive loops, so the exponent and coefficient
// derivatives of a bra primitive and the pair's bra-centre gradient are
// summed over the segment by lane exchanges before one atomic; with one
// charge per lane (at most 64 charges) each charge's gradient is summed over
// the pair's primitives in registers and over the wavefront's segments
// before one atomic per charge and wavefront.

// value of the primitive integral with bra powers ia, ket powers jb
template <int LI, int LJ>
struct DerivTables {
    static constexpr int IM = LI + 3, JS = LJ + 3, JV = LJ + 1, NR = (LI + 2 + LJ) / 2 + 1;
    double s[3][IM][JS];   // 1D overlaps (S, T)
    double t[3][IM][JV];   // 1D kinetic factors (T)
    double I[NR][3][IM][JV];  // 1D Rys integrals times weights folded later (V)
    double w[NR];
};

template <int LI, int LJ, int TYP>
__device__ __forceinline__ double deriv_val(const DerivTables<LI, LJ> &T, const int *ia, const int *jb)
{
    if constexpr (TYP == 0) return T.s[0][ia[0]][jb[0]] * T.s[1][ia[1]][jb[1]] * T.s[2][ia[2]][jb[2]];
    if constexpr (TYP == 1) {
        const double sx = T.s[0][ia[0]][jb[0]], sy = T.s[1][ia[1]][jb[1]], sz = T.s[2][ia[2]][jb[2]];
        return T.t[0][ia[0]][jb[0]] * sy * sz + sx * T.t[1][ia[1]][jb[1]] * sz + sx * sy * T.t[2][ia[2]][jb[2]];
    }
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < DerivTables<LI, LJ>::NR; r++)
        acc += T.w[r] * T.I[r][0][ia[0]][jb[0]] * T.I[r][1][ia[1]][jb[1]] * T.I[r][2][ia[2]][jb[2]];
    return acc;
}

// W-weighted sums of one charge (V) or of S/T over the components of the
// pair: adds exponent / coefficient derivatives of bra primitive ip and the
// bra-centre gradient; dS mode writes the matrix entries instead.  The
// component loops are unrolled up to d shells, so that the tables are
// indexed with constants and stay in registers.
template <int LI, int LJ, int TYP>
__device__ __forceinline__ void deriv_accumulate(const DerivTables<LI, LJ> &T, double pref, double alpha, double ca,
                                                 const Shell1Dev &A, const Shell1Dev &B, int ip,
                                                 const double *W, int nao, double &dex, double &dco, double gA[3],
                                                 double *dS, long long s1, long long L)
{
    constexpr int NFI = (LI + 1) * (LI + 2) / 2, NFJ = (LJ + 1) * (LJ + 2) / 2;
    constexpr int UNR = (LI <= 2 && LJ <= 2) ? 64 : 1;
    const size_t n = (size_t)nao;
#pragma unroll UNR
    for (int fj = 0; fj < NFJ; fj++)
#pragma unroll UNR
        for (int fi = 0; fi < NFI; fi++) {
            const int ia[3] = {c1x(LI, fi), c1y(LI, fi), LI - c1x(LI, fi) - c1y(LI, fi)};
            const int jb[3] = {c1x(LJ, fj), c1y(LJ, fj), LJ - c1x(LJ, fj) - c1y(LJ, fj)};
            const double v0 = pref * deriv_val<LI, LJ, TYP>(T, ia, jb);
            double de = 0.0, dA[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                int up[3] = {ia[0], ia[1], ia[2]};
                up[d] += 2;
                de -= pref * deriv_val<LI, LJ, TYP>(T, up, jb);
                int p1[3] = {ia[0], ia[1], ia[2]};
                p1[d] += 1;
                double g = 2.0 * alpha * pref * deriv_val<LI, LJ, TYP>(T, p1, jb);
                if (ia[d] > 0) {
                    int m1[3] = {ia[0], ia[1], ia[2]};
                    m1[d] -= 1;
                    g -= ia[d] * pref * deriv_val<LI, LJ, TYP>(T, m1, jb);
                }
                dA[d] = g;
            }
            const size_t a = A.ao + fi, b = B.ao + fj;
            if (W) {
                const double w2 = 2.0 * W[a * n + b];
                dex += w2 * ca * de;
                dco += w2 * v0;
#pragma unroll
                for (int d = 0; d < 3; d++) gA[d] += w2 * ca * dA[d];
            }
            if (TYP == 0 && dS) {
                const long long ie = A.pexp + ip, ic = A.pcoef + ip;
                if (ie >= s1 && ie < s1 + L) {
                    atomicAdd(&dS[(a * n + b) * L + (ie - s1)], ca * de);
                    atomicAdd(&dS[(b * n + a) * L + (ie - s1)], ca * de);
                }
                if (ic >= s1 && ic < s1 + L) {
                    atomicAdd(&dS[(a * n + b) * L + (ic - s1)], v0);
                    atomicAdd(&dS[(b * n + a) * L + (ic - s1)], v0);
                }
            }
        }
}

// sum over the aligned segment of cw lanes (a power of two; every lane of
// the segment active)
__device__ __forceinline__ double seg_sum(double v, int cw)
{
    for (int off = 1; off < cw; off <<= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int LI, int LJ, int TYP>
__global__ __launch_bounds__(64) void int1e_deriv_kernel(const int2 *__restrict__ pairs, int npair,
                                                         const Shell1Dev *__restrict__ sh,
                                                         const double *__restrict__ env,
                                                         const double4 *__restrict__ nuc, int nnuc, int cwlog, int nao,
                                                         const double *__restrict__ W, double *__restrict__ gatom,
                                                         double *__restrict__ denv, long long env_l,
                                                         double *__restrict__ dS, long long s1, long long L)
{
    using TT = DerivTables<LI, LJ>;
    // lanes per pair: 1 for S and T, 2^cwlog for V
    const int cw = TYP == 2 ? 1 << cwlog : 1;
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long t = TYP == 2 ? gid >> cwlog : gid;
    const int cs = TYP == 2 ? (int)(gid & (cw - 1)) : 0;
    if (TYP != 2 && t >= npair) return;
    // V: every lane stays to the end (the wavefront-wide sum of the charge
    // gradients); a segment past the last pair runs no primitive loop
    const bool live = t < npair;
    const bool one = nnuc <= cw;  // at most one charge per lane
    const int2 pr = pairs[live ? t : 0];
    const Shell1Dev A = sh[pr.x], B = sh[pr.y];  // A.pad = atom of A
    double AB[3], AB2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        AB[d] = A.r[d] - B.r[d];
        AB2 += AB[d] * AB[d];
    }
    const double cf = fac_sp_dev(LI) * fac_sp_dev(LJ);
    TT T;
    double gAt[3] = {0.0, 0.0, 0.0}, gC[3] = {0.0, 0.0, 0.0};
    const int nip = live ? A.nprim : 0;
    for (int ip = 0; ip < nip; ip++) {
        const double a = env[A.pexp + ip], ca = env[A.pcoef + ip];
        double dex = 0.0, dco = 0.0;
        for (int jp = 0; jp < B.nprim; jp++) {
            const double b = env[B.pexp + jp], cb = env[B.pcoef + jp];
            const double p = a + b, h = 0.5 / p;
            double P[3];
#pragma unroll
            for (int d = 0; d < 3; d++) P[d] = (a * A.r[d] + b * B.r[d]) / p;
            const double pref0 = cb * cf * exp(-a * b / p * AB2);  // bra coefficient left out
            if constexpr (TYP <= 1) {
                const double s00 = sqrt(M_PI / p);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double PA = P[d] - A.r[d], PB = P[d] - B.r[d];
#pragma unroll
                    for (int i = 0; i < TT::IM; i++)
#pragma unroll
                        for (int j = 0; j < TT::JS; j++) {
                            double v;
                            if (i == 0 && j == 0) v = s00;
                            else if (i > 0) {
                                v = PA * T.s[d][i - 1][j];
                                if (i > 1) v += (i - 1) * h * T.s[d][i - 2][j];
                                if (j > 0) v += j * h * T.s[d][i - 1][j - 1];
                            } else {
                                v = PB * T.s[d][i][j - 1];
                                if (j > 1) v += (j - 1) * h * T.s[d][i][j - 2];
                            }
                            T.s[d][i][j] = v;
                        }
                    if constexpr (TYP == 1) {
#pragma unroll
                        for (int i = 0; i < TT::IM; i++)
#pragma unroll
                            for (int j = 0; j < TT::JV; j++) {
                                double tt = -2.0 * b * (2 * j + 1) * T.s[d][i][j] + 4.0 * b * b * T.s[d][i][j + 2];
                                if (j >= 2) tt += j * (j - 1) * T.s[d][i][j - 2];
                                T.t[d][i][j] = -0.5 * tt;
                            }
                    }
                }
                deriv_accumulate<LI, LJ, TYP>(T, pref0, a, ca, A, B, ip, W, nao, dex, dco, gAt, dS, s1, L);
            } else {
                for (int c = cs; c < nnuc; c += cw) {
                    const double4 C = nuc[c];
                    const double PC[3] = {P[0] - C.x, P[1] - C.y, P[2] - C.z};
                    const double PC2 = PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2];
                    double rv[TT::NR];
                    rys_nodes<TT::NR>(p * PC2, rv, T.w);
                    constexpr int LT = LI + 2 + LJ;
#pragma unroll
                    for (int r = 0; r < TT::NR; r++) {
                        const double b10 = (1.0 - rv[r]) * h;
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            const double c00 = (P[d] - A.r[d]) - rv[r] * PC[d];
                            double H[LT + 1][LJ + 1];
                            H[0][0] = 1.0;
                            if (LT > 0) H[1][0] = c00;
#pragma unroll
                            for (int i = 1; i < LT; i++) H[i + 1][0] = c00 * H[i][0] + i * b10 * H[i - 1][0];
#pragma unroll
                            for (int j = 1; j <= LJ; j++)
#pragma unroll
                                for (int i = 0; i <= LT - j; i++) H[i][j] = H[i + 1][j - 1] + AB[d] * H[i][j - 1];
#pragma unroll
                            for (int i = 0; i < TT::IM; i++)
#pragma unroll
                                for (int j = 0; j < TT::JV; j++) T.I[r][d][i][j] = H[i][j];
                        }
                    }
                    double gA[3] = {0.0, 0.0, 0.0};
                    const double pref = -C.w * 2.0 * M_PI / p * pref0;
                    deriv_accumulate<LI, LJ, 2>(T, pref, a, ca, A, B, ip, W, nao, dex, dco, gA, nullptr, 0, 0);
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        gAt[d] += gA[d];
                        if (one) gC[d] += gA[d];
                        else if (gatom) atomicAdd(&gatom[3 * c + d], -gA[d]);
                    }
                }
            }
        }
        if (TYP == 2) {
            // the pair's lanes run the same primitive loops: all active here
            dex = seg_sum(dex, cw);
            dco = seg_sum(dco, cw);
        }
        if (denv && cs == 0) {
            if (A.pexp + ip < env_l) atomicAdd(&denv[A.pexp + ip], dex);
            if (A.pcoef + ip < env_l) atomicAdd(&denv[A.pcoef + ip], dco);
        }
    }
    if (!gatom) return;
    if constexpr (TYP == 2) {
        if (live) {
#pragma unroll
            for (int d = 0; d < 3; d++) gAt[d] = seg_sum(gAt[d], cw);
        }
        if (one) {
            // every lane of the wavefront: charge cs summed over the segments
#pragma unroll
            for (int d = 0; d < 3; d++)
                for (int off = cw; off < 64; off <<= 1) gC[d] += __shfl_xor(gC[d], off, 64);
            if ((int)threadIdx.x < cw && cs < nnuc) {
#pragma unroll
                for (int d = 0; d < 3; d++) atomicAdd(&gatom[3 * cs + d], -gC[d]);
            }
        }
    }
    if (live && cs == 0) {
#pragma unroll
        for (int d = 0; d < 3; d++) atomicAdd(&gatom[3 * A.pad + d], gAt[d]);
    }
}

hipError_t launch_int1e_deriv(int li, int lj, const int2 *pairs, int npair, const Shell1Dev *sh, const double *env,
                              const double4 *nuc, int nnuc, int nao, int typ, const double *W, double *gatom,
                              double *denv, long long env_l, double *dS, long long s1, long long L, hipStream_t st)
{
    if (npair <= 0) return hipSuccess;
    if (typ < 0 || typ > 2 || (dS && typ != 0)) return hipErrorInvalidValue;
    if (typ == 2 && nnuc <= 0) return hipSuccess;
    // V: lanes per pair, the charge count rounded up to a power of two, at
    // most 64 (LIBRINT_V_LANES lowers the cap: fewer lanes per pair, each
    // taking several charges)
    int cwmax = 6;
    if (const char *e = getenv("LIBRINT_V_LANES")) {
        const int v = atoi(e);
        cwmax = 0;
        while (cwmax < 6 && (2 << cwmax) <= v) cwmax++;
    }
    int cwlog = 0;
    if (typ == 2)
        while ((1 << cwlog) < nnuc && cwlog < cwmax) cwlog++;
    const long long nthr = (long long)npair << (typ == 2 ? cwlog : 0);
    if ((nthr + 63) / 64 > 0x7fffffffLL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((nthr + 63) / 64));
#define X(a, b)                                                                                                        \
    if (li == a && lj == b) {                                                                                          \
        if (typ == 0)                                                                                                  \
            hipLaunchKernelGGL((int1e_deriv_kernel<a, b, 0>), grid, dim3(64), 0, st, pairs, npair, sh, env, nuc, nnuc, \
                               cwlog, nao, W, gatom, denv, env_l, dS, s1, L);                                          \
        else if (typ == 1)                                                                                             \
            hipLaunchKernelGGL((int1e_deriv_kernel<a, b, 1>), grid, dim3(64), 0, st, pairs, npair, sh, env, nuc, nnuc, \
                               cwlog, nao, W, gatom, denv, env_l, dS, s1, L);                                          \
        else                                                                                                           \
            hipLaunchKernelGGL((int1e_deriv_kernel<a, b, 2>), grid, dim3(64), 0, st, pairs, npair, sh, env, nuc, nnuc, \
                               cwlog, nao, W, gatom, denv, env_l, dS, s1, L);                                          \
        return hipGetLastError();                                                                                      \
    }
    X(0, 0) X(0, 1) X(0, 2) X(1, 0) X(1, 1) X(1, 2) X(2, 0) X(2, 1) X(2, 2)
    X(0, 3) X(1, 3) X(2, 3) X(3, 0) X(3, 1) X(3, 2) X(3, 3)
    X(0, 4) X(1, 4) X(2, 4) X(3, 4) X(4, 0) X(4, 1) X(4, 2) X(4, 3) X(4, 4)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_int1e(int li, int lj, const int2 *pairs, int npair, const Shell1Dev *sh, const double *env,
                        const double4 *nuc, int nnuc, int nao, double *S, double *T, double *V, hipStream_t st)
{
    if (npair <= 0) return hipSuccess;
    const int mask = (S ? 1 : 0) | (T ? 2 : 0) | (V ? 4 : 0);
    const dim3 grid((npair + 127) / 128);
#define X(a, b)                                                                                                        \
    if (li == a && lj == b) {                                                                                          \
        hipLaunchKernelGGL((int1e_kernel<a, b>), grid, dim3(128), 0, st, pairs, npair, sh, env, nuc, nnuc, nao, mask, \
                           S, T, V);                                                                                   \
        return hipGetLastError();                                                                                      \
    }
    X(0, 0) X(0, 1) X(0, 2) X(1, 0) X(1, 1) X(1, 2) X(2, 0) X(2, 1) X(2, 2)
    X(0, 3) X(1, 3) X(2, 3) X(3, 0) X(3, 1) X(3, 2) X(3, 3)
    X(0, 4) X(1, 4) X(2, 4) X(3, 4) X(4, 0) X(4, 1) X(4, 2) X(4, 3) X(4, 4)
#undef X
    return hipErrorInvalidValue;
}

} // namespace librint
