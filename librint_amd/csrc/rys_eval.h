// Rys quadrature nodes and weights from the piecewise-Chebyshev table
// (rys_table.inc, produced by gen_rys_table.py).
//
// Output convention: v_r = t_r^2 in [0, 1) (not the u = v/(1-v) of
// CINTrys_roots, reference src/rys_roots.rs:3036-3064; the recurrence
// coefficients below are written in v directly) and plain weights with
// sum_r w_r = F_0(x).
//
// x < RYS_XCUT[n]: segment k = floor(x) of unit length, local variable
//   s = 2(x - k) - 1 in [-1, 1], 2n polynomials of degree RYS_DEG = 11 in s
//   (monomial coefficients RYS_CHEB[off + (k*(DEG+1) + c)*2n + j], c the
//   power) evaluated with Estrin's scheme.
// x >= RYS_XCUT[n]: Gauss-Hermite limit v_r = h_r / x, w_r = H_r / sqrt(x).
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RYS_HD __host__ __device__
#else
#include <cmath>
#define RYS_HD
#endif

namespace librint {

#if defined(__HIP_DEVICE_COMPILE__)
#define RYS_STORAGE __device__ static const
#else
#define RYS_STORAGE static const
#endif
#include "rys_table.inc"
#include "rys_table_f32.inc"
#undef RYS_STORAGE

// Estrin evaluation of a degree-11 polynomial sum_c a[c*stride] s^c
// (depth 4 instead of the 11 dependent steps of Horner's rule).
RYS_HD inline double rys_estrin11(const double *a, int stride, double s, double s2, double s4, double s8)
{
    const double q0 = fma(a[1 * stride], s, a[0]);
    const double q1 = fma(a[3 * stride], s, a[2 * stride]);
    const double q2 = fma(a[5 * stride], s, a[4 * stride]);
    const double q3 = fma(a[7 * stride], s, a[6 * stride]);
    const double q4 = fma(a[9 * stride], s, a[8 * stride]);
    const double q5 = fma(a[11 * stride], s, a[10 * stride]);
    const double r0 = fma(q1, s2, q0);
    const double r1 = fma(q3, s2, q2);
    const double r2 = fma(q5, s2, q4);
    return fma(r2, s8, fma(r1, s4, r0));
}

template <int N>
RYS_HD inline void rys_nodes(double x, double *v, double *w)
{
    static_assert(N >= 1 && N <= RYS_NMAX, "root count");
    constexpr int NV = 2 * N;
    if (x >= (double)RYS_XCUT[N]) {
#if defined(__HIP_DEVICE_COMPILE__)
        double rs = __builtin_amdgcn_rsq(x);
        double e = fma(-x * rs, rs, 1.0);
        rs = fma(0.5 * rs, e, rs);
        e = fma(-x * rs, rs, 1.0);
        rs = fma(0.5 * rs, e, rs);
        const double rx = rs * rs;
#else
        const double rx = 1.0 / x;
        const double rs = sqrt(rx);
#endif
#pragma unroll
        for (int r = 0; r < N; r++) {
            v[r] = RYS_HERM_NODE2[RYS_HOFF[N] + r] * rx;
            w[r] = RYS_HERM_W[RYS_HOFF[N] + r] * rs;
        }
        return;
    }
    static_assert(RYS_DEG == 11, "Estrin evaluation is written for degree 11");
    const int k = (int)x;
    const double s = 2.0 * (x - (double)k) - 1.0;
    const double s2 = s * s, s4 = s2 * s2, s8 = s4 * s4;
    const double *c = RYS_CHEB + RYS_OFF[N] + k * (RYS_DEG + 1) * NV;
#pragma unroll
    for (int j = 0; j < N; j++) {
        v[j] = rys_estrin11(c + j, NV, s, s2, s4, s8);
        w[j] = rys_estrin11(c + N + j, NV, s, s2, s4, s8);
    }
}

// Runtime-n dispatch (host code and tests).
RYS_HD inline void rys_nodes_n(int n, double x, double *v, double *w)
{
    switch (n) {
    case 1: rys_nodes<1>(x, v, w); break;
    case 2: rys_nodes<2>(x, v, w); break;
    case 3: rys_nodes<3>(x, v, w); break;
    case 4: rys_nodes<4>(x, v, w); break;
    case 5: rys_nodes<5>(x, v, w); break;
    case 6: rys_nodes<6>(x, v, w); break;
    case 7: rys_nodes<7>(x, v, w); break;
    case 8: rys_nodes<8>(x, v, w); break;
    default: rys_nodes<9>(x, v, w); break;
    }
}

} // namespace librint
