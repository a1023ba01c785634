// Rys-quadrature ERI core for one contracted shell quartet, fully unrolled for
// a compile-time angular class (LA LB | LC LD).
//
// Convention (same recurrence as reference src/eri.rs:263-399, the lj form of
// libcint's CINTg0_2e_2d + CINTg0_lj2d_4d, src/g2e.rs:316-606):
//   * the 2D VRR builds momentum on centre b (bra) and d (ket) up to
//     NMAX = LA+LB and MMAX = LC+LD;
//   * the HRR moves it b -> a and d -> c.
// Pairs are stored with LA <= LB and LC <= LD, so the HRR always moves the
// smaller momentum.
//
// Per-root g buffers, index  i + DK*k + DL*l + DJ*j  (i on a, k on c, l on d,
// j on b), DK = LA+1, DL = DK*(LC+1), DJ = DL*(MMAX+1).
//
// The contracted block is accumulated into eri[NF] with
//   f = fa + NFA*(fb + NFB*(fc + NFC*fd))   (a fastest, cint2e_cart order).
#pragma once

#include "rys_eval.h"

namespace librint {

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// cartesian powers of component idx of shell l, x-major descending
// (reference src/cint_bas.rs:234-255)
__host__ __device__ constexpr int cart_x(int l, int idx)
{
    int n = 0;
    for (int lx = l; lx >= 0; lx--)
        for (int ly = l - lx; ly >= 0; ly--) {
            if (n == idx) return lx;
            n++;
        }
    return 0;
}
__host__ __device__ constexpr int cart_y(int l, int idx)
{
    int n = 0;
    for (int lx = l; lx >= 0; lx--)
        for (int ly = l - lx; ly >= 0; ly--) {
            if (n == idx) return ly;
            n++;
        }
    return 0;
}
__host__ __device__ constexpr int cart_z(int l, int idx) { return l - cart_x(l, idx) - cart_y(l, idx); }

template <int LA, int LB, int LC, int LD>
struct Class {
    static constexpr int NMAX = LA + LB;
    static constexpr int MMAX = LC + LD;
    static constexpr int NROOTS = (NMAX + MMAX) / 2 + 1;
    static constexpr int DK = LA + 1;
    static constexpr int DL = DK * (LC + 1);
    static constexpr int DJ = DL * (MMAX + 1);
    static constexpr int GS = DJ * (NMAX + 1);
    static constexpr int NFA = ncart(LA), NFB = ncart(LB), NFC = ncart(LC), NFD = ncart(LD);
    static constexpr int NF = NFA * NFB * NFC * NFD;
};

// 2D VRR on one axis (reference src/eri.rs:318-370)
template <int NMAX, int MMAX, int DL, int DJ, typename T>
__device__ __forceinline__ void vrr_axis(T *g, T cn, T cm, T b10, T b01, T b00)
{
    if constexpr (NMAX > 0) {
        g[DJ] = cn * g[0];
#pragma unroll
        for (int n = 1; n < NMAX; n++) g[(n + 1) * DJ] = cn * g[n * DJ] + (T)n * b10 * g[(n - 1) * DJ];
    }
    if constexpr (MMAX > 0) {
        g[DL] = cm * g[0];
#pragma unroll
        for (int m = 1; m < MMAX; m++) g[(m + 1) * DL] = cm * g[m * DL] + (T)m * b01 * g[(m - 1) * DL];
        if constexpr (NMAX > 0) {
            g[DJ + DL] = cm * g[DJ] + b00 * g[0];
#pragma unroll
            for (int m = 1; m < MMAX; m++)
                g[DJ + (m + 1) * DL] = cm * g[DJ + m * DL] + (T)m * b01 * g[DJ + (m - 1) * DL] + b00 * g[m * DL];
        }
    }
    if constexpr (NMAX > 1) {
#pragma unroll
        for (int m = 1; m <= MMAX; m++) {
#pragma unroll
            for (int n = 1; n < NMAX; n++)
                g[m * DL + (n + 1) * DJ] =
                    cn * g[m * DL + n * DJ] + (T)n * b10 * g[m * DL + (n - 1) * DJ] + (T)m * b00 * g[(m - 1) * DL + n * DJ];
        }
    }
}

// HRR on one axis: b -> a with rab = Rb - Ra, d -> c with rcd = Rd - Rc
// (reference src/eri.rs:372-398)
template <int LA, int LB, int LC, int NMAX, int MMAX, int DK, int DL, int DJ, typename T>
__device__ __forceinline__ void hrr_axis(T *g, T rab, T rcd)
{
#pragma unroll
    for (int i = 1; i <= LA; i++)
#pragma unroll
        for (int j = 0; j <= NMAX - i; j++)
#pragma unroll
            for (int l = 0; l <= MMAX; l++) {
                const int p = j * DJ + l * DL + i;
                g[p] = rab * g[p - 1] + g[p - 1 + DJ];
            }
#pragma unroll
    for (int j = 0; j <= LB; j++)
#pragma unroll
        for (int k = 1; k <= LC; k++)
#pragma unroll
            for (int l = 0; l <= MMAX - k; l++) {
                const int p = j * DJ + l * DL + k * DK;
#pragma unroll
                for (int n = 0; n < DK; n++) g[p + n] = rcd * g[p + n - DK] + g[p + n - DK + DL];
            }
}

// A chunk of the contracted block: the outputs whose component index on
// centre SD (0 a, 1 b, 2 c, 3 d) equals S0, or the whole block for SD < 0.
// Classes whose block does not fit a lane's registers evaluate it chunk by
// chunk (the primitive loop repeated per chunk), so that only one chunk's
// outputs are live at a time.  Chunk-local index
//   f = (fa - A0) + NA ((fb - B0) + NB ((fc - C0) + NC (fd - D0))).
template <int LA, int LB, int LC, int LD, int SD, int S0>
struct Chunk {
    using C = Class<LA, LB, LC, LD>;
    static constexpr int A0 = SD == 0 ? S0 : 0, NA = SD == 0 ? 1 : C::NFA;
    static constexpr int B0 = SD == 1 ? S0 : 0, NB = SD == 1 ? 1 : C::NFB;
    static constexpr int C0 = SD == 2 ? S0 : 0, NC = SD == 2 ? 1 : C::NFC;
    static constexpr int D0 = SD == 3 ? S0 : 0, ND = SD == 3 ? 1 : C::NFD;
    static constexpr int NFL = NA * NB * NC * ND;
    __host__ __device__ static constexpr int idx(int fa, int fb, int fc, int fd)
    {
        return (fa - A0) + NA * ((fb - B0) + NB * ((fc - C0) + NC * (fd - D0)));
    }
};

// Accumulate sum_r gx*gy*gz of one root into the contracted block (or the
// chunk (SD, S0) of it).
template <int LA, int LB, int LC, int LD, typename T, typename A, int SD = -1, int S0 = 0>
__device__ __forceinline__ void gout_acc(A *eri, const T *gx, const T *gy, const T *gz)
{
    using C = Class<LA, LB, LC, LD>;
    using K = Chunk<LA, LB, LC, LD, SD, S0>;
#pragma unroll
    for (int fd = K::D0; fd < K::D0 + K::ND; fd++)
#pragma unroll
        for (int fc = K::C0; fc < K::C0 + K::NC; fc++)
#pragma unroll
            for (int fb = K::B0; fb < K::B0 + K::NB; fb++)
#pragma unroll
                for (int fa = K::A0; fa < K::A0 + K::NA; fa++) {
                    const int ix = cart_x(LB, fb) * C::DJ + cart_x(LD, fd) * C::DL + cart_x(LC, fc) * C::DK + cart_x(LA, fa);
                    const int iy = cart_y(LB, fb) * C::DJ + cart_y(LD, fd) * C::DL + cart_y(LC, fc) * C::DK + cart_y(LA, fa);
                    const int iz = cart_z(LB, fb) * C::DJ + cart_z(LD, fd) * C::DL + cart_z(LC, fc) * C::DK + cart_z(LA, fa);
                    eri[K::idx(fa, fb, fc, fd)] += (A)(gx[ix] * gy[iy] * gz[iz]);
                }
}

// Primitive-pair record (one per surviving primitive pair of a shell pair).
//   e  = ai + aj
//   K  = sqrt(2 pi^{5/2}) ca cb fac_sp(la) fac_sp(lb) exp(-ai aj/e |Ra-Rb|^2) / e
//        so that the quartet prefactor 2 pi^{5/2} K_ij K_kl / (e_ij e_kl
//        sqrt(e_ij + e_kl)) of reference src/g2e.rs:5547-5560 is K_p K_q rsqrt(e_p + e_q)
//   h  = 1 / (2 e)
//   pb = P - Rb  (P the Gaussian product centre, Rb the VRR centre).
struct PrimPair {
    double e;
    double K;
    double h;
    double pbx, pby, pbz;
    double pad0, pad1;
};

// Reciprocal and reciprocal square root for positive normal arguments:
// hardware estimate plus two Newton steps (relative error ~1 ulp).
// 1/sqrt(x) for x > 0: the hardware estimate (relative error about 5e-8)
// refined by one third-order step, y (1 + e/2 + 3e^2/8) with e = 1 - x y^2;
// the dropped term 5e^3/16 is below 1e-21, so the result is within about
// one ulp, like two Newton steps but with three fewer operations.
__device__ __forceinline__ double rsqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y, e * fma(0.375, e, 0.5), y);
}
__device__ __forceinline__ double rcp_pos(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// Reciprocal square root in the working precision.
__device__ __forceinline__ double rsqrt_t(double x) { return rsqrt_pos(x); }
__device__ __forceinline__ float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }

// Rys table block of one root count staged in LDS, in the working precision
// T: double (rys_table.inc, degree 11) or float (rys_table_f32.inc, degree 5).
// rt = [seg][c][2N] (v then w), or for WONLY (all-s quartets, weights only)
// [seg][c]; segments STRIDE elements apart, STRIDE padded so that it is 16
// bytes past a multiple of 128 bytes: lanes of a wavefront read the same
// coefficient row of different segments, and this keeps consecutive segments
// in different LDS bank groups.
template <typename T, int N, bool WONLY>
struct RysTab {
    static constexpr int NV = WONLY ? 1 : 2 * N;
    static constexpr int DEG = sizeof(T) == 8 ? RYS_DEG : RYS32_DEG;
    static constexpr int NCOEF = (DEG + 1) * NV;
    static constexpr int EPB = 128 / (int)sizeof(T);  // elements per 128 bytes
    static constexpr int PAD16 = 16 / (int)sizeof(T); // elements per 16 bytes
    static constexpr int STRIDE = NCOEF + ((PAD16 - NCOEF % EPB) % EPB + EPB) % EPB;
    static constexpr int LEN = RYS_NSEG[N] * STRIDE;
};

// copy the block of root count N into LDS (all threads of the block take part)
template <typename T, int N, bool WONLY>
__device__ __forceinline__ void rys_stage(T *rt)
{
    using R = RysTab<T, N, WONLY>;
    for (int i = threadIdx.x; i < R::LEN; i += blockDim.x) {
        const int seg = i / R::STRIDE, c = i - seg * R::STRIDE;
        const int g = seg * R::NCOEF + c;  // index in the unpadded (weights-only) block
        T v = (T)0;
        if (c < R::NCOEF) {
            if constexpr (sizeof(T) == 8)
                v = R::NV == 2 * N ? (T)RYS_CHEB[RYS_OFF[N] + g] : (T)RYS_CHEB[RYS_OFF[N] + 2 * g + 1];
            else
                v = R::NV == 2 * N ? (T)RYS32_TAB[RYS32_OFF[N] + g] : (T)RYS32_TAB[RYS32_OFF[N] + 2 * g + 1];
        }
        rt[i] = v;
    }
}

template <typename T>
__device__ __forceinline__ T estrin_t(const T *a, int st, T s, T s2, T s4, T s8)
{
    if constexpr (sizeof(T) == 8) {
        return rys_estrin11(a, st, s, s2, s4, s8);
    } else {
        const T q0 = fma(a[st], s, a[0]);
        const T q1 = fma(a[3 * st], s, a[2 * st]);
        const T q2 = fma(a[5 * st], s, a[4 * st]);
        return fma(q2, s4, fma(q1, s2, q0));
    }
}

template <typename T, int N, bool WONLY>
__device__ __forceinline__ void rys_nodes_lds(T x, const T *rt, T *v, T *w)
{
    using R = RysTab<T, N, WONLY>;
    if (x >= (T)RYS_NSEG[N]) {
        // large x: the Gauss-Hermite limit, evaluated in FP64 in both
        // working precisions (FP32 keeps only the rounded nodes and weights)
        const double rs = rsqrt_pos((double)x);
        const double rx = rs * rs;
#pragma unroll
        for (int r = 0; r < N; r++) {
            v[r] = WONLY ? (T)0 : (T)(RYS_HERM_NODE2[RYS_HOFF[N] + r] * rx);
            w[r] = (T)(RYS_HERM_W[RYS_HOFF[N] + r] * rs);
        }
        return;
    }
    const int k = (int)x;
    const T s = (T)2 * (x - (T)k) - (T)1;
    const T s2 = s * s, s4 = s2 * s2, s8 = s4 * s4;
    const T *c = rt + k * R::STRIDE;
    if constexpr (WONLY) {
        w[0] = estrin_t<T>(c, 1, s, s2, s4, s8);
        v[0] = (T)0;
    } else {
#pragma unroll
        for (int j = 0; j < N; j++) {
            v[j] = estrin_t<T>(c + j, R::NV, s, s2, s4, s8);
            w[j] = estrin_t<T>(c + N + j, R::NV, s, s2, s4, s8);
        }
    }
}

template <int LA, int LB, int LC, int LD, typename T = double>
struct RysLds {
    using C = Class<LA, LB, LC, LD>;
    static constexpr int N = C::NROOTS;
    static constexpr bool WONLY = (C::NMAX + C::MMAX) == 0;
    using Tab = RysTab<T, N, WONLY>;
    static constexpr int LEN = Tab::LEN;
};

// One primitive quartet, all roots, in working precision T with the Rys
// table block in LDS; accumulated into eri.  qd = (Q - Rd) - (Rb - Rd) =
// Q - Rb of the ket primitive pair (P - Q = (P - Rb) - qd), rab = Rb - Ra,
// rcd = Rd - Rc.
// Contribution of one Rys root (node v = t^2, weight w) of one primitive
// quartet: recurrence coefficients, VRR, HRR and the gout products.
template <int LA, int LB, int LC, int LD, typename T, int SD = -1, int S0 = 0>
__device__ __forceinline__ void root_term(T *eri, const PrimPair &p, const PrimPair &q, const T pq[3], T aij, T akl,
                                          T rs, T fac1, T vr, T wr, const T rab[3], const T rcd[3])
{
    using C = Class<LA, LB, LC, LD>;
    const T tr = vr * rs;
    const T b00 = (T)0.5 * tr;
    const T b10 = (T)p.h * ((T)1 - akl * tr);
    const T b01 = (T)q.h * ((T)1 - aij * tr);
    const T ta = akl * tr;
    const T tc = aij * tr;
    T gx[C::GS], gy[C::GS], gz[C::GS];
    gx[0] = (T)1;
    gy[0] = (T)1;
    gz[0] = wr * fac1;
    vrr_axis<C::NMAX, C::MMAX, C::DL, C::DJ>(gx, (T)p.pbx - ta * pq[0], (T)q.pbx + tc * pq[0], b10, b01, b00);
    vrr_axis<C::NMAX, C::MMAX, C::DL, C::DJ>(gy, (T)p.pby - ta * pq[1], (T)q.pby + tc * pq[1], b10, b01, b00);
    vrr_axis<C::NMAX, C::MMAX, C::DL, C::DJ>(gz, (T)p.pbz - ta * pq[2], (T)q.pbz + tc * pq[2], b10, b01, b00);
    if constexpr (LA > 0 || LC > 0) {
        hrr_axis<LA, LB, LC, C::NMAX, C::MMAX, C::DK, C::DL, C::DJ>(gx, rab[0], rcd[0]);
        hrr_axis<LA, LB, LC, C::NMAX, C::MMAX, C::DK, C::DL, C::DJ>(gy, rab[1], rcd[1]);
        hrr_axis<LA, LB, LC, C::NMAX, C::MMAX, C::DK, C::DL, C::DJ>(gz, rab[2], rcd[2]);
    }
    gout_acc<LA, LB, LC, LD, T, T, SD, S0>(eri, gx, gy, gz);
}

// One primitive quartet, all roots, in working precision T with the Rys
// table block in LDS; accumulated into eri.  qd = (Q - Rd) - (Rb - Rd) =
// Q - Rb of the ket primitive pair (P - Q = (P - Rb) - qd), rab = Rb - Ra,
// rcd = Rd - Rc.
template <int LA, int LB, int LC, int LD, typename T, int SD = -1, int S0 = 0>
__device__ __forceinline__ void prim_quartet_lds(T *eri, const PrimPair &p, const PrimPair &q, const T qd[3],
                                                 const T rab[3], const T rcd[3], const T *rt)
{
    using C = Class<LA, LB, LC, LD>;
    constexpr int N = C::NROOTS;
    const T aij = (T)p.e, akl = (T)q.e;
    const T rsq = rsqrt_t(aij + akl);
    const T rs = rsq * rsq;
    const T a0 = aij * akl * rs;
    const T pq[3] = {(T)p.pbx - qd[0], (T)p.pby - qd[1], (T)p.pbz - qd[2]};
    const T rr = pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2];
    const T fac1 = (T)p.K * (T)q.K * rsq;
    T v[N], w[N];
    rys_nodes_lds<T, N, RysLds<LA, LB, LC, LD, T>::WONLY>(a0 * rr, rt, v, w);
#pragma unroll
    for (int r = 0; r < N; r++)
        root_term<LA, LB, LC, LD, T, SD, S0>(eri, p, q, pq, aij, akl, rs, fac1, v[r], w[r], rab, rcd);
}

// One primitive quartet, all roots, accumulated into eri (double precision).
// rbd = Rb - Rd, rab = Rb - Ra, rcd = Rd - Rc.
// One primitive quartet with the Rys table read from global memory (double).
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void prim_quartet_fp64(double *eri, const PrimPair &p, const PrimPair &q, const double rbd[3],
                                                  const double rab[3], const double rcd[3])
{
    using C = Class<LA, LB, LC, LD>;
    const double aij = p.e, akl = q.e;
    const double rsq = rsqrt_pos(aij + akl);
    const double rs = rsq * rsq;
    const double a0 = aij * akl * rs;
    const double pq[3] = {p.pbx - q.pbx + rbd[0], p.pby - q.pby + rbd[1], p.pbz - q.pbz + rbd[2]};
    const double rr = pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2];
    const double fac1 = p.K * q.K * rsq;
    double v[C::NROOTS], w[C::NROOTS];
    rys_nodes<C::NROOTS>(a0 * rr, v, w);
#pragma unroll
    for (int r = 0; r < C::NROOTS; r++) {
        // v = t^2;  b00 = t^2/(2(aij+akl)), b10 = (1 - akl t^2/(aij+akl))/(2 aij), ...
        const double t2 = v[r];
        const double tr = t2 * rs;
        const double b00 = 0.5 * tr;
        const double b10 = p.h * (1.0 - akl * tr);
        const double b01 = q.h * (1.0 - aij * tr);
        const double ta = akl * tr; // weight of (P-Q) in the bra shift
        const double tc = aij * tr;
        double gx[C::GS], gy[C::GS], gz[C::GS];
        gx[0] = 1.0;
        gy[0] = 1.0;
        gz[0] = w[r] * fac1;
        vrr_axis<C::NMAX, C::MMAX, C::DL, C::DJ>(gx, p.pbx - ta * pq[0], q.pbx + tc * pq[0], b10, b01, b00);
        vrr_axis<C::NMAX, C::MMAX, C::DL, C::DJ>(gy, p.pby - ta * pq[1], q.pby + tc * pq[1], b10, b01, b00);
        vrr_axis<C::NMAX, C::MMAX, C::DL, C::DJ>(gz, p.pbz - ta * pq[2], q.pbz + tc * pq[2], b10, b01, b00);
        if constexpr (LA > 0 || LC > 0) {
            hrr_axis<LA, LB, LC, C::NMAX, C::MMAX, C::DK, C::DL, C::DJ>(gx, rab[0], rcd[0]);
            hrr_axis<LA, LB, LC, C::NMAX, C::MMAX, C::DK, C::DL, C::DJ>(gy, rab[1], rcd[1]);
            hrr_axis<LA, LB, LC, C::NMAX, C::MMAX, C::DK, C::DL, C::DJ>(gz, rab[2], rcd[2]);
        }
        gout_acc<LA, LB, LC, LD>(eri, gx, gy, gz);
    }
}

// ---- chains in LDS (team kernels: one lane per (root, axis) chain) ----

__device__ __forceinline__ double sel3(int ax, double x, double y, double z) { return ax == 0 ? x : (ax == 1 ? y : z); }

// 2D VRR of one chain in LDS, rolled over m so that the compiler keeps only a
// few values live (the fully unrolled form forwards every store to its loads
// and spills).  g(n, m) at n*DJ + m*DL.
template <int NMAX, int MMAX, int DL, int DJ>
__device__ __forceinline__ void vrr_rolled(double *g, double cn, double cm, double b10, double b01, double b00)
{
    if constexpr (NMAX > 0) {
        double gm1 = 0.0, g0 = g[0];
#pragma unroll
        for (int n = 0; n < NMAX; n++) {
            const double g1 = cn * g0 + (double)n * b10 * gm1;
            g[(n + 1) * DJ] = g1;
            gm1 = g0;
            g0 = g1;
        }
    }
#pragma unroll 1
    for (int m = 0; m < MMAX; m++) {
        const double mb01 = (double)m * b01;
        double prev = 0.0;  // g(n-1, m)
#pragma unroll
        for (int n = 0; n <= NMAX; n++) {
            const double cur = g[n * DJ + m * DL];
            double v = cm * cur;
            if (m > 0) v += mb01 * g[n * DJ + (m - 1) * DL];
            if (n > 0) v += (double)n * b00 * prev;
            g[n * DJ + (m + 1) * DL] = v;
            prev = cur;
        }
    }
}

// HRR of one chain in LDS: b -> a with rab = Rb - Ra (k = 0 plane), then
// d -> c with rcd = Rd - Rc, rolled over the outer indices.
template <int LA, int LB, int LC, int NMAX, int MMAX, int DK, int DL, int DJ>
__device__ __forceinline__ void hrr_rolled(double *g, double rab, double rcd)
{
#pragma unroll 1
    for (int i = 1; i <= LA; i++)
#pragma unroll 1
        for (int j = 0; j <= NMAX - i; j++)
#pragma unroll
            for (int l = 0; l <= MMAX; l++) {
                const int p = j * DJ + l * DL + i;
                g[p] = rab * g[p - 1] + g[p - 1 + DJ];
            }
#pragma unroll 1
    for (int j = 0; j <= LB; j++)
#pragma unroll 1
        for (int k = 1; k <= LC; k++)
#pragma unroll 1
            for (int l = 0; l <= MMAX - k; l++) {
                const int p = j * DJ + l * DL + k * DK;
#pragma unroll
                for (int n = 0; n < DK; n++) g[p + n] = rcd * g[p + n - DK] + g[p + n - DK + DL];
            }
}

// Node and weight of root r only (each chain lane needs one root; evaluating
// all 2N polynomials per lane would keep 24 coefficient rows live).
template <int N>
__device__ __forceinline__ void rys_one(double x, int r, double &v, double &w)
{
    if (x >= (double)RYS_XCUT[N]) {
        const double rs = rsqrt_pos(x);
        v = RYS_HERM_NODE2[RYS_HOFF[N] + r] * (rs * rs);
        w = RYS_HERM_W[RYS_HOFF[N] + r] * rs;
        return;
    }
    const int k = (int)x;
    const double s = 2.0 * (x - (double)k) - 1.0;
    const double s2 = s * s, s4 = s2 * s2, s8 = s4 * s4;
    const double *c = RYS_CHEB + RYS_OFF[N] + k * (RYS_DEG + 1) * 2 * N;
    v = rys_estrin11(c + r, 2 * N, s, s2, s4, s8);
    w = rys_estrin11(c + N + r, 2 * N, s, s2, s4, s8);
}

} // namespace librint
