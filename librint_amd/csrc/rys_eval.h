// Rys quadrature nodes and weights from the piecewise-Chebyshev table
// (rys_table.inc, produced by gen_rys_table.py).
//
// Output convention: v_r = t_r^2 in [0, 1) (not the u = v/(1-v) of
// CINTrys_roots, reference src/rys_roots.rs:3036-3064; the recurrence
// coefficients below are written in v directly) and plain weights with
// sum_r w_r = F_0(x).
//
// x < RYS_XCUT[n]: segment k = floor(x) of unit length, local variable
//   s = 2(x - k) - 1 in [-1, 1], 2n Clenshaw recurrences of degree RYS_DEG
//   over contiguous coefficients RYS_CHEB[off + (k*(DEG+1) + c)*2n + j].
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
#undef RYS_STORAGE

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
    const int k = (int)x;
    const double s = 2.0 * (x - (double)k) - 1.0;
    const double s2 = 2.0 * s;
    const double *c = RYS_CHEB + RYS_OFF[N] + k * (RYS_DEG + 1) * NV;
    double b1[NV], b2[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) {
        b1[j] = 0.0;
        b2[j] = 0.0;
    }
#pragma unroll
    for (int m = RYS_DEG; m >= 1; m--) {
#pragma unroll
        for (int j = 0; j < NV; j++) {
            const double b0 = fma(s2, b1[j], c[m * NV + j] - b2[j]);
            b2[j] = b1[j];
            b1[j] = b0;
        }
    }
#pragma unroll
    for (int j = 0; j < N; j++) {
        v[j] = fma(s, b1[j], c[j] - b2[j]);
        w[j] = fma(s, b1[N + j], c[N + j] - b2[N + j]);
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
