// Adjoint of the two-electron energy E2 = 1/2 tr(P G[P]) at fixed P:
//   dE2/dtheta = sum over canonical quartets and functions f = (a b|c d) of
//                W_f d(ab|cd)_f / dtheta,
//   W_f = s (4 P_ab P_cd - P_ac P_bd - P_ad P_bc),  s = 1/2^{[a=b]+[c=d]+[ab=cd]},
// which is the seed of reference src/dscf.rs:390-470 (1/2 (P_ab P_cd -
// 1/2 P_ac P_bd) summed over the distinct index images of a canonical
// quartet).  theta runs over the nuclear coordinates and over every
// primitive exponent and contraction coefficient (the env entries of the
// reference's dR_c, src/dscf.rs:517-524).
//
// Derivatives of a primitive cartesian Gaussian x_A^a exp(-alpha x_A^2):
//   d/dA_x      = 2 alpha G(a + 1_x) - a_x G(a - 1_x)
//   d/dalpha    = -(G(a + 2_x) + G(a + 2_y) + G(a + 2_z))
//   d/dc        = (primitive part of the contracted integral) / c
// The Rys 1D integrals are built with the momentum of the differentiated
// pair (SIDE 0: bra centres a, b; SIDE 1: ket centres c, d, computed as the
// bra of (cd|ab)) raised by 2; every derivative swaps one axis factor of
// sum_r Ix Iy Iz for its shifted combination.
//
// Work mapping: a team of TS lanes owns one quartet; lanes c < 3*NROOTS
// build the (root, axis) chains of the raised recurrence in LDS, then every
// lane contracts its share of the weighted outputs.  The 64/TS teams of a
// wavefront walk the window pairs of one task (one bin: same class and
// primitive counts, so all loops are uniform) while the task's fixed pair is
// the differentiated one.  Derivatives go to per-shell-primitive slots (not
// to env entries shared by every atom of an element); each task adds its
// wavefront-reduced sums with one atomic per slot and value.
#include <hip/hip_runtime.h>

#include "eri_core.h"
#include "grad.h"

namespace librint {

__device__ __constant__ signed char GCART[3][6][3] = {
    {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}}};

__host__ __device__ constexpr int pow2_at_least(int x)
{
    int p = 1;
    while (p < x) p *= 2;
    return p;
}
__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }

// Shape of the raised recurrence (g index i + DK k + DL l + DJ j as in the
// Fock kernels, with the raised momenta).
template <int LA, int LB, int LC, int LD, int SIDE>
struct GShape {
    static constexpr int EA = SIDE == 0 ? 2 : 0, EC = SIDE == 1 ? 2 : 0;
    static constexpr int LA2 = LA + EA, LB2 = LB + EA, LC2 = LC + EC;
    static constexpr int NMAX = LA + LB + EA, MMAX = LC + LD + EC;
    static constexpr int NROOTS = (LA + LB + LC + LD + 2) / 2 + 1;
    static constexpr int DK = LA2 + 1;
    static constexpr int DL = DK * (LC2 + 1);
    static constexpr int DJ = DL * (MMAX + 1);
    static constexpr int GS = DJ * (NMAX + 1);
    static constexpr int NFA = ncart(LA), NFB = ncart(LB), NFC = ncart(LC), NFD = ncart(LD);
    static constexpr int NF = NFA * NFB * NFC * NFD;
    static constexpr int NCH = 3 * NROOTS;
    // team: enough lanes for the chains and at most ~8 outputs per lane
    static constexpr int TS = cmin(64, cmax(pow2_at_least(NCH), pow2_at_least((NF + 7) / 8)));
    static constexpr int NT = 64 / TS;  // teams per wavefront
    static constexpr int NPL = (NF + TS - 1) / TS;
    // shift strides of the two differentiated centres: a (i) and b (j), or c (k) and d (l)
    static constexpr int ST0 = SIDE == 0 ? 1 : DK, ST1 = SIDE == 0 ? DJ : DL;
    static_assert(GS < 1024, "packed g indices use 10 bits");
};

template <int TS>
__device__ __forceinline__ double team_sum(double v)
{
#pragma unroll
    for (int off = TS / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}


struct GPair {
    int sa, sb, off, npp, aoa, aob;
    double rb[3], rab[3];
};

__device__ __forceinline__ GPair gload_pair(const DevBasis &B, int p)
{
    GPair v;
    const int4 pi = B.pair_info[p];
    v.sa = pi.x;
    v.sb = pi.y;
    v.off = pi.z;
    v.npp = pi.w;
    v.aoa = B.ao_off[v.sa];
    v.aob = B.ao_off[v.sb];
    const double4 g0 = B.pair_rb[p];
    const double4 g1 = B.pair_rab[p];
    v.rb[0] = g0.x;
    v.rb[1] = g0.y;
    v.rb[2] = g0.z;
    v.rab[0] = g1.x;
    v.rab[1] = g1.y;
    v.rab[2] = g1.z;
    return v;
}

// One (root, axis) chain c = 3 r + axis of the raised recurrence for one
// primitive quartet.  rbd = Rb - Rd, rab = Rb - Ra, rcd = Rd - Rc.
template <int LA, int LB, int LC, int LD, int SIDE>
__device__ __forceinline__ void grad_chain(double *g, int c, const PrimPair &p, const PrimPair &q, const double rbd[3],
                                           const double rab[3], const double rcd[3])
{
    using S = GShape<LA, LB, LC, LD, SIDE>;
    const double aij = p.e, akl = q.e;
    const double rsq = rsqrt_pos(aij + akl);
    const double rs = rsq * rsq;
    const double a0 = aij * akl * rs;
    const double pqx = p.pbx - q.pbx + rbd[0], pqy = p.pby - q.pby + rbd[1], pqz = p.pbz - q.pbz + rbd[2];
    const double rr = pqx * pqx + pqy * pqy + pqz * pqz;
    const double fac1 = p.K * q.K * rsq;
    const int r = c / 3, ax = c - 3 * r;
    double t2, wr;
    rys_one<S::NROOTS>(a0 * rr, r, t2, wr);
    const double tr = t2 * rs;
    const double b00 = 0.5 * tr;
    const double b10 = p.h * (1.0 - akl * tr);
    const double b01 = q.h * (1.0 - aij * tr);
    const double pq = sel3(ax, pqx, pqy, pqz);
    const double cn = sel3(ax, p.pbx, p.pby, p.pbz) - akl * tr * pq;
    const double cm = sel3(ax, q.pbx, q.pby, q.pbz) + aij * tr * pq;
    double *gc = g + c * S::GS;
    if constexpr (S::GS <= 64) {
        // short chains: the recurrences unrolled in registers, then one
        // store of the chain to LDS
        double gl[S::GS];
#pragma unroll
        for (int e = 0; e < S::GS; e++) gl[e] = 0.0;
        gl[0] = (ax == 2) ? wr * fac1 : 1.0;
        vrr_axis<S::NMAX, S::MMAX, S::DL, S::DJ>(gl, cn, cm, b10, b01, b00);
        if constexpr (S::LA2 > 0 || S::LC2 > 0)
            hrr_axis<S::LA2, S::LB2, S::LC2, S::NMAX, S::MMAX, S::DK, S::DL, S::DJ>(
                gl, sel3(ax, rab[0], rab[1], rab[2]), sel3(ax, rcd[0], rcd[1], rcd[2]));
#pragma unroll
        for (int e = 0; e < S::GS; e++) gc[e] = gl[e];
    } else {
        gc[0] = (ax == 2) ? wr * fac1 : 1.0;
        vrr_rolled<S::NMAX, S::MMAX, S::DL, S::DJ>(gc, cn, cm, b10, b01, b00);
        if constexpr (S::LA2 > 0 || S::LC2 > 0)
            hrr_rolled<S::LA2, S::LB2, S::LC2, S::NMAX, S::MMAX, S::DK, S::DL, S::DJ>(
                gc, sel3(ax, rab[0], rab[1], rab[2]), sel3(ax, rcd[0], rcd[1], rcd[2]));
    }
}

template <int LA, int LB, int LC, int LD, int SIDE>
__global__ __launch_bounds__(64) void grad_kernel(const JKTask *__restrict__ tasks, int ntask, DevBasis B,
                                                  const int2 *__restrict__ pp_idx, const int2 *__restrict__ shell_info,
                                                  const double *__restrict__ coef, const double *__restrict__ D, int nao,
                                                  double *__restrict__ gatom, double *__restrict__ dprim)
{
    // The differentiated pair X (bra for SIDE 0, ket for SIDE 1) always takes
    // the bra role of the raised recurrence: the ket side is computed as the
    // bra side of (cd|ab), which has the same energy weights, so both passes
    // raise the momentum along the unrolled n direction of the VRR.
    constexpr int XA = SIDE == 0 ? LA : LC, XB = SIDE == 0 ? LB : LD;
    constexpr int YA = SIDE == 0 ? LC : LA, YB = SIDE == 0 ? LD : LB;
    using S = GShape<XA, XB, YA, YB, 0>;
    using C = Class<LA, LB, LC, LD>;
    __shared__ double g[S::NT * S::NCH * S::GS];
    const int tid = blockIdx.x;
    if (tid >= ntask) return;
    const JKTask t = tasks[tid];
    const int lane = threadIdx.x;
    const int team = lane / S::TS, tl = lane - team * S::TS;
    double *gt = g + team * S::NCH * S::GS;
    // The task's fixed pair is the differentiated one (SIDE 0: the canonical
    // bra; SIDE 1: the canonical ket, from a ket-major task list whose window
    // runs over canonical bras), so every team of the wavefront adds to the
    // same derivative slots and the sums are reduced in registers over the
    // whole task before one set of atomics.
    const GPair sd = gload_pair(B, t.bra);
    const int2 si0 = shell_info[sd.sa], si1 = shell_info[sd.sb];
    // SIDE 0 runs over the Fock build's task list, whose heavy quartets may
    // be split by bra primitive pairs; the ket-major list is never split
    int o_begin = 0, o_end = sd.npp;
    if (SIDE == 0) task_prims(t, sd.npp, o_begin, o_end);
    const size_t n = (size_t)nao;
    const int niter = (t.nket + S::NT - 1) / S::NT;
    // nuclear derivatives of the two centres (per-lane partial sums); the
    // primitive sums of record o live in lane o (o < 64) or lane o - 64
    double gx0[3] = {0.0, 0.0, 0.0}, gx1[3] = {0.0, 0.0, 0.0};
    double pe0[2] = {0.0, 0.0}, pe1[2] = {0.0, 0.0}, pv[2] = {0.0, 0.0};
    for (int it = 0; it < niter; it++) {
        const int kk = it * S::NT + team;
        const bool valid = kk < t.nket;
        const int otp = t.ket0 + (valid ? kk : 0);
        const GPair ot = gload_pair(B, otp);
        const GPair &br = SIDE == 0 ? sd : ot;  // canonical bra (a b|
        const GPair &kt = SIDE == 0 ? ot : sd;  // canonical ket |c d)
        double s = valid ? 1.0 : 0.0;
        if (br.sa == br.sb) s *= 0.5;
        if (kt.sa == kt.sb) s *= 0.5;
        if (t.bra == otp) s *= 0.5;
        // this lane's outputs: energy weight, packed g indices, powers of the
        // two differentiated centres (2 bits per axis)
        double W[S::NPL];
        int gi[S::NPL], pw[S::NPL];
#pragma unroll
        for (int k = 0; k < S::NPL; k++) {
            const int f = tl + S::TS * k;
            W[k] = 0.0;
            gi[k] = 0;
            pw[k] = 0;
            if (f < S::NF) {
                const int fa = f % C::NFA, t1 = f / C::NFA;
                const int fb = t1 % C::NFB, t2 = t1 / C::NFB;
                const int fc = t2 % C::NFC, fd = t2 / C::NFC;
                const size_t a = br.aoa + fa, b = br.aob + fb, c = kt.aoa + fc, d = kt.aob + fd;
                W[k] = s * (4.0 * D[a * n + b] * D[c * n + d] - D[a * n + c] * D[b * n + d] - D[a * n + d] * D[b * n + c]);
                int packed = 0, pws = 0;
#pragma unroll
                for (int x = 0; x < 3; x++) {
                    const int ia = GCART[LA][fa][x], ib = GCART[LB][fb][x], ic = GCART[LC][fc][x], id = GCART[LD][fd][x];
                    const int xa = SIDE == 0 ? ia : ic, xb = SIDE == 0 ? ib : id;
                    const int ya = SIDE == 0 ? ic : ia, yb = SIDE == 0 ? id : ib;
                    packed |= (xa + ya * S::DK + yb * S::DL + xb * S::DJ) << (10 * x);
                    pws |= (xa << (2 * x)) | (xb << (6 + 2 * x));
                }
                gi[k] = packed;
                pw[k] = pws;
            }
        }
        const double rxy[3] = {sd.rb[0] - ot.rb[0], sd.rb[1] - ot.rb[1], sd.rb[2] - ot.rb[2]};
        for (int o = o_begin; o < o_end; o++) {
            const PrimPair po = B.pp[sd.off + o];
            const double al0 = po.pad0, al1 = po.pad1;  // exponents of the two centres
            double e0 = 0.0, e1 = 0.0, val = 0.0;
            for (int i = 0; i < ot.npp; i++) {
                const PrimPair pi = B.pp[ot.off + i];
                if (tl < S::NCH) grad_chain<XA, XB, YA, YB, 0>(gt, tl, po, pi, rxy, sd.rab, ot.rab);
                __syncthreads();
#pragma unroll
                for (int k = 0; k < S::NPL; k++) {
                    if (tl + S::TS * k >= S::NF) continue;
                    const int b[3] = {gi[k] & 1023, (gi[k] >> 10) & 1023, (gi[k] >> 20) & 1023};
                    const double Wk = W[k];
#pragma unroll 1
                    for (int r = 0; r < S::NROOTS; r++) {
                        const double *G0 = gt + (3 * r) * S::GS, *G1 = G0 + S::GS, *G2 = G1 + S::GS;
                        const double I[3] = {G0[b[0]], G1[b[1]], G2[b[2]]};
                        val += Wk * I[0] * I[1] * I[2];
#pragma unroll
                        for (int x = 0; x < 3; x++) {
                            const double *Gx = x == 0 ? G0 : (x == 1 ? G1 : G2);
                            const double o1 = Wk * I[(x + 1) % 3] * I[(x + 2) % 3];
                            const int n0 = (pw[k] >> (2 * x)) & 3, n1 = (pw[k] >> (6 + 2 * x)) & 3;
                            double d0 = 2.0 * al0 * Gx[b[x] + S::ST0];
                            if (n0) d0 -= n0 * Gx[b[x] - S::ST0];
                            double d1 = 2.0 * al1 * Gx[b[x] + S::ST1];
                            if (n1) d1 -= n1 * Gx[b[x] - S::ST1];
                            gx0[x] += d0 * o1;
                            gx1[x] += d1 * o1;
                            e0 -= Gx[b[x] + 2 * S::ST0] * o1;
                            e1 -= Gx[b[x] + 2 * S::ST1] * o1;
                        }
                    }
                }
                __syncthreads();
            }
            e0 = team_sum<64>(e0);
            e1 = team_sum<64>(e1);
            val = team_sum<64>(val);
            if (o < 64) {
                if (lane == o) {
                    pe0[0] += e0;
                    pe1[0] += e1;
                    pv[0] += val;
                }
            } else if (o < 128) {
                if (lane == o - 64) {
                    pe0[1] += e0;
                    pe1[1] += e1;
                    pv[1] += val;
                }
            } else if (lane == 0) {
                const int2 ij = pp_idx[sd.off + o];
                const int k0 = si0.y + ij.x, k1 = si1.y + ij.y;
                atomicAdd(&dprim[2 * k0], e0);
                atomicAdd(&dprim[2 * k1], e1);
                atomicAdd(&dprim[2 * k0 + 1], val / coef[k0]);
                atomicAdd(&dprim[2 * k1 + 1], val / coef[k1]);
            }
        }
    }
#pragma unroll
    for (int x = 0; x < 3; x++) {
        const double v0 = team_sum<64>(gx0[x]);
        const double v1 = team_sum<64>(gx1[x]);
        if (lane == 0) {
            atomicAdd(&gatom[3 * si0.x + x], v0);
            atomicAdd(&gatom[3 * si1.x + x], v1);
        }
    }
#pragma unroll
    for (int slot = 0; slot < 2; slot++) {
        const int o = 64 * slot + lane;
        if (o < sd.npp) {
            const int2 ij = pp_idx[sd.off + o];
            const int k0 = si0.y + ij.x, k1 = si1.y + ij.y;
            atomicAdd(&dprim[2 * k0], pe0[slot]);
            atomicAdd(&dprim[2 * k1], pe1[slot]);
            atomicAdd(&dprim[2 * k0 + 1], pv[slot] / coef[k0]);
            atomicAdd(&dprim[2 * k1 + 1], pv[slot] / coef[k1]);
        }
    }
}

#define LR_GRAD_CLASSES(X)                                                                                             \
    X(0, 0, 0, 0) X(0, 1, 0, 0) X(0, 1, 0, 1) X(1, 1, 0, 0) X(1, 1, 0, 1) X(1, 1, 1, 1) X(0, 2, 0, 0) X(0, 2, 0, 1)   \
        X(0, 2, 1, 1) X(0, 2, 0, 2) X(1, 2, 0, 0) X(1, 2, 0, 1) X(1, 2, 1, 1) X(1, 2, 0, 2) X(1, 2, 1, 2)             \
            X(2, 2, 0, 0) X(2, 2, 0, 1) X(2, 2, 1, 1) X(2, 2, 0, 2) X(2, 2, 1, 2) X(2, 2, 2, 2)

hipError_t launch_grad(int la, int lb, int lc, int ld, int side, const JKTask *tasks, int ntask, const DevBasis &B,
                       const int2 *pp_idx, const int2 *shell_info, const double *coef, const double *D, int nao,
                       double *gatom, double *dprim, hipStream_t st)
{
    if (ntask <= 0) return hipSuccess;
    if (side != 0 && side != 1) return hipErrorInvalidValue;
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d) {                                                                    \
        if (side == 0)                                                                                                 \
            hipLaunchKernelGGL((grad_kernel<a, b, c, d, 0>), dim3(ntask), dim3(64), 0, st, tasks, ntask, B, pp_idx,    \
                               shell_info, coef, D, nao, gatom, dprim);                                                \
        else                                                                                                           \
            hipLaunchKernelGGL((grad_kernel<a, b, c, d, 1>), dim3(ntask), dim3(64), 0, st, tasks, ntask, B, pp_idx,    \
                               shell_info, coef, D, nao, gatom, dprim);                                                \
        return hipGetLastError();                                                                                      \
    }
    LR_GRAD_CLASSES(X)
#undef X
    return hipErrorInvalidValue;
}

} // namespace librint
