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

#include "int1e_gpu.h"
#include "rys_eval.h"

namespace librint {

__device__ __constant__ signed char C1E[5][15][3] = {
    {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 1, 1}, {1, 0, 2}, {0, 3, 0}, {0, 2, 1}, {0, 1, 2}, {0, 0, 3}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{4, 0, 0}, {3, 1, 0}, {3, 0, 1}, {2, 2, 0}, {2, 1, 1}, {2, 0, 2}, {1, 3, 0}, {1, 2, 1}, {1, 1, 2}, {1, 0, 3}, {0, 4, 0}, {0, 3, 1}, {0, 2, 2}, {0, 1, 3}, {0, 0, 4}}};

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
                            const int i = C1E[LI][fi][d], j = C1E[LJ][fj][d];
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
                                bv[fi + NFI * fj] += fw * I[0][C1E[LI][fi][0]][C1E[LJ][fj][0]] *
                                                     I[1][C1E[LI][fi][1]][C1E[LJ][fj][1]] *
                                                     I[2][C1E[LI][fi][2]][C1E[LJ][fj][2]];
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
