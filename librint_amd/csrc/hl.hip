// Two-electron integrals of the angular classes with an f or g shell
// (l = 3, 4; up to 9 Rys roots, g arrays of up to 2025 entries per chain),
// for the Fock build, the Schwarz diagonal, the tensor and shell blocks.
// Semantics as the l <= 2 kernels: reference src/eri.rs:98-399,409-692 for
// one contracted quartet, src/scf.rs:316-410 for its use in the Fock build.
//
// Work mapping: one workgroup of one wavefront per quartet (kets of a task
// in turn).  The class shape is a kernel argument, so one kernel serves all
// 99 such classes.  For each primitive quartet the (root, axis) chains of
// the 2D recurrence and the horizontal transfer are built in LDS by lanes
// c < 3*rb, rb roots at a time (rb from the LDS budget: 3 roots of (ff|ff),
// 1 of (gg|gg)); then every lane adds sum_r gx gy gz of its outputs
// f = lane + 64k to the quartet's contracted block, which lives in a global
// scratch slice of the workgroup (up to 50625 doubles, too many for
// registers or LDS).  The block then feeds the J/K update (LDS rows flushed
// with global atomics), the Schwarz maximum, the tensor scatter or the copy.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "eri_core.h"
#include "hl.h"

namespace librint {

// cartesian component powers in libcint order, l <= 4
__device__ __constant__ signed char HCART[5][15][3] = {
    {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 1, 1}, {1, 0, 2}, {0, 3, 0}, {0, 2, 1}, {0, 1, 2}, {0, 0, 3}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{4, 0, 0}, {3, 1, 0}, {3, 0, 1}, {2, 2, 0}, {2, 1, 1}, {2, 0, 2}, {1, 3, 0}, {1, 2, 1}, {1, 1, 2}, {1, 0, 3}, {0, 4, 0}, {0, 3, 1}, {0, 2, 2}, {0, 1, 3}, {0, 0, 4}}};

// LDS budget of the chains (bytes): 3 chains of one root of (gg|gg) fit
static constexpr int HL_CHAIN_LDS = 48 * 1024 + 1024;
// J/K rows: six blocks of at most 15 x 15 entries
static constexpr int HL_JK_ROWS = 6 * 225;
static constexpr int HL_MAX_GROUPS = 8192;
// LDS budget of the adjoint's raised chains (bytes): one root of the
// raised (gg|gg) chains (3 x 3465 doubles) fits
static constexpr int HL_GRAD_CHAIN_LDS = 96 * 1024;

__device__ __forceinline__ double wave_sum64(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

HLShape hl_shape(int la, int lb, int lc, int ld)
{
    HLShape S;
    S.la = la;
    S.lb = lb;
    S.lc = lc;
    S.ld = ld;
    S.nfa = ncart(la);
    S.nfb = ncart(lb);
    S.nfc = ncart(lc);
    S.nfd = ncart(ld);
    S.nf = S.nfa * S.nfb * S.nfc * S.nfd;
    S.nmax = la + lb;
    S.mmax = lc + ld;
    S.nroots = (S.nmax + S.mmax) / 2 + 1;
    S.dk = la + 1;
    S.dl = S.dk * (lc + 1);
    S.dj = S.dl * (S.mmax + 1);
    S.gs = S.dj * (S.nmax + 1);
    const int per_root = 3 * S.gs * (int)sizeof(double);
    S.rb = std::max(1, std::min(S.nroots, HL_CHAIN_LDS / per_root));
    return S;
}

// Node v = t^2 and weight of root r of the n-root rule (run-time n <= 9).
__device__ __forceinline__ void rys_one_rt(int n, double x, int r, double &v, double &w)
{
    if (x >= (double)RYS_XCUT[n]) {
        const double rs = rsqrt_pos(x);
        v = RYS_HERM_NODE2[RYS_HOFF[n] + r] * (rs * rs);
        w = RYS_HERM_W[RYS_HOFF[n] + r] * rs;
        return;
    }
    const int k = (int)x;
    const double s = 2.0 * (x - (double)k) - 1.0;
    const double s2 = s * s, s4 = s2 * s2, s8 = s4 * s4;
    const double *c = RYS_CHEB + RYS_OFF[n] + k * (RYS_DEG + 1) * 2 * n;
    v = rys_estrin11(c + r, 2 * n, s, s2, s4, s8);
    w = rys_estrin11(c + n + r, 2 * n, s, s2, s4, s8);
}

// 2D VRR of one chain, g(n, m) at n*dj + m*dl (reference src/eri.rs:318-370).
__device__ __forceinline__ void vrr_rt(double *g, const HLShape &S, double cn, double cm, double b10, double b01,
                                       double b00)
{
    double gm1 = 0.0, g0 = g[0];
    for (int n = 0; n < S.nmax; n++) {
        const double g1 = cn * g0 + (double)n * b10 * gm1;
        g[(n + 1) * S.dj] = g1;
        gm1 = g0;
        g0 = g1;
    }
    for (int m = 0; m < S.mmax; m++) {
        const double mb01 = (double)m * b01;
        double prev = 0.0;  // g(n-1, m)
        for (int n = 0; n <= S.nmax; n++) {
            const double cur = g[n * S.dj + m * S.dl];
            double v = cm * cur;
            if (m > 0) v += mb01 * g[n * S.dj + (m - 1) * S.dl];
            if (n > 0) v += (double)n * b00 * prev;
            g[n * S.dj + (m + 1) * S.dl] = v;
            prev = cur;
        }
    }
}

// HRR of one chain: b -> a with rab = Rb - Ra on the k = 0 plane, then
// d -> c with rcd = Rd - Rc (reference src/eri.rs:371-399).
__device__ __forceinline__ void hrr_rt(double *g, const HLShape &S, double rab, double rcd)
{
    for (int i = 1; i <= S.la; i++)
        for (int j = 0; j <= S.nmax - i; j++)
            for (int l = 0; l <= S.mmax; l++) {
                const int p = j * S.dj + l * S.dl + i;
                g[p] = rab * g[p - 1] + g[p - 1 + S.dj];
            }
    for (int j = 0; j <= S.lb; j++)
        for (int k = 1; k <= S.lc; k++)
            for (int l = 0; l <= S.mmax - k; l++) {
                const int p = j * S.dj + l * S.dl + k * S.dk;
                for (int n = 0; n < S.dk; n++) g[p + n] = rcd * g[p + n - S.dk] + g[p + n - S.dk + S.dl];
            }
}

// output f of the block (a fastest) -> components and g offsets
struct HLOut {
    int fa, fb, fc, fd;
    int ix, iy, iz;
};

__device__ __forceinline__ HLOut hl_out(const HLShape &S, int f)
{
    HLOut o;
    o.fa = f % S.nfa;
    const int t1 = f / S.nfa;
    o.fb = t1 % S.nfb;
    const int t2 = t1 / S.nfb;
    o.fc = t2 % S.nfc;
    o.fd = t2 / S.nfc;
    int idx[3];
#pragma unroll
    for (int x = 0; x < 3; x++)
        idx[x] = HCART[S.lb][o.fb][x] * S.dj + HCART[S.ld][o.fd][x] * S.dl + HCART[S.lc][o.fc][x] * S.dk +
                 HCART[S.la][o.fa][x];
    o.ix = idx[0];
    o.iy = idx[1];
    o.iz = idx[2];
    return o;
}

// Contracted block of the quartet (bra|ket) into scr[0, nf): each lane owns
// the entries f = lane + 64k.
__device__ void hl_block(const HLShape &S, const PairView &bra, const PairView &ket, const DevBasis &B, double *g,
                         double *scr, int lane)
{
    for (int f = lane; f < S.nf; f += 64) scr[f] = 0.0;
    const double rbd[3] = {bra.rb[0] - ket.rb[0], bra.rb[1] - ket.rb[1], bra.rb[2] - ket.rb[2]};
    for (int ib = 0; ib < bra.npp; ib++) {
        const PrimPair p = B.pp[bra.off + ib];
        for (int ik = 0; ik < ket.npp; ik++) {
            const PrimPair q = B.pp[ket.off + ik];
            const double aij = p.e, akl = q.e;
            const double rsq = rsqrt_pos(aij + akl);
            const double rs = rsq * rsq;
            const double a0 = aij * akl * rs;
            const double pq[3] = {p.pbx - q.pbx + rbd[0], p.pby - q.pby + rbd[1], p.pbz - q.pbz + rbd[2]};
            const double rr = pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2];
            const double fac1 = p.K * q.K * rsq;
            const double x = a0 * rr;
            for (int r0 = 0; r0 < S.nroots; r0 += S.rb) {
                const int nb = min(S.rb, S.nroots - r0);
                if (lane < 3 * nb) {
                    const int r = r0 + lane / 3, ax = lane - 3 * (lane / 3);
                    double t2, wr;
                    rys_one_rt(S.nroots, x, r, t2, wr);
                    const double tr = t2 * rs;
                    const double b00 = 0.5 * tr;
                    const double b10 = p.h * (1.0 - akl * tr);
                    const double b01 = q.h * (1.0 - aij * tr);
                    const double pqa = sel3(ax, pq[0], pq[1], pq[2]);
                    const double cn = sel3(ax, p.pbx, p.pby, p.pbz) - akl * tr * pqa;
                    const double cm = sel3(ax, q.pbx, q.pby, q.pbz) + aij * tr * pqa;
                    double *gc = g + lane * S.gs;
                    gc[0] = (ax == 2) ? wr * fac1 : 1.0;
                    vrr_rt(gc, S, cn, cm, b10, b01, b00);
                    hrr_rt(gc, S, sel3(ax, bra.rab[0], bra.rab[1], bra.rab[2]),
                           sel3(ax, ket.rab[0], ket.rab[1], ket.rab[2]));
                }
                __syncthreads();
                for (int f = lane; f < S.nf; f += 64) {
                    const HLOut o = hl_out(S, f);
                    double v = 0.0;
                    for (int c = 0; c < nb; c++) {
                        const double *G = g + 3 * c * S.gs;
                        v += G[o.ix] * G[S.gs + o.iy] * G[2 * S.gs + o.iz];
                    }
                    scr[f] += v;
                }
                __syncthreads();
            }
        }
    }
}

// MODE 0..2: list modes of launch_list; MODE 3: Fock build over JKTasks.
template <int MODE>
__global__ __launch_bounds__(64) void hl_kernel(HLShape S, const JKTask *__restrict__ jtasks,
                                                const ListTask *__restrict__ ltasks, int ntask,
                                                const unsigned char *__restrict__ flags, int accept, DevBasis B,
                                                const double *__restrict__ D, double *__restrict__ M,
                                                double *__restrict__ out, int nao, double *__restrict__ scr_all)
{
    extern __shared__ double g[];
    __shared__ double rows[HL_JK_ROWS];
    __shared__ double red[64];
    const int lane = threadIdx.x;
    double *scr = scr_all + (size_t)blockIdx.x * S.nf;
    const size_t n = (size_t)nao;
    const int nab = S.nfa * S.nfb, ncd = S.nfc * S.nfd, nac = S.nfa * S.nfc, nbc = S.nfb * S.nfc,
              nad = S.nfa * S.nfd, nbd = S.nfb * S.nfd;
    double *mab = rows, *mcd = rows + nab, *mac = mcd + ncd, *mbc = mac + nac, *mad = mbc + nbc, *mbd = mad + nad;
    const int nket_rows = ncd + nac + nbc + nad + nbd;
    for (int ti = blockIdx.x; ti < ntask; ti += gridDim.x) {
        int brap, ket0, nket;
        long long oo = 0;
        if constexpr (MODE == 3) {
            if (flags && !((accept >> flags[ti]) & 1)) continue;
            const JKTask t = jtasks[ti];
            brap = t.bra;
            ket0 = t.ket0;
            nket = t.nket;
        } else {
            const ListTask t = ltasks[ti];
            brap = t.bra;
            ket0 = t.ket;
            nket = 1;
            oo = t.out;
        }
        const PairView bra = load_pair(B, brap);
        if constexpr (MODE == 3) {
            for (int e = lane; e < nab; e += 64) mab[e] = 0.0;
        }
        for (int kk = 0; kk < nket; kk++) {
            const int ketp = ket0 + kk;
            const PairView ket = load_pair(B, ketp);
            hl_block(S, bra, ket, B, g, scr, lane);
            if constexpr (MODE == 0) {
                for (int f = lane; f < S.nf; f += 64) out[oo + f] = scr[f];
            } else if constexpr (MODE == 1) {
                double m = 0.0;
                for (int f = lane; f < S.nf; f += 64) {
                    const HLOut o = hl_out(S, f);
                    if (o.fa == o.fc && o.fb == o.fd) m = fmax(m, fabs(scr[f]));
                }
                red[lane] = m;
                __syncthreads();
                if (lane == 0) {
                    double mm = 0.0;
                    for (int e = 0; e < 64; e++) mm = fmax(mm, red[e]);
                    out[oo] = sqrt(mm);
                }
                __syncthreads();
            } else if constexpr (MODE == 2) {
                // entries related by a same-shell pair or by bra = ket map to
                // the same 8 addresses: only the canonical one writes, so every
                // image carries one value
                const bool sab = bra.sa == bra.sb, scd = ket.sa == ket.sb, sbk = brap == ketp;
                for (int f = lane; f < S.nf; f += 64) {
                    const HLOut o = hl_out(S, f);
                    if ((sab && o.fa < o.fb) || (scd && o.fc < o.fd) ||
                        (sbk && o.fa * S.nfb + o.fb < o.fc * S.nfd + o.fd))
                        continue;
                    const double v = scr[f];
                    const size_t a = bra.aoa + o.fa, b = bra.aob + o.fb, c = ket.aoa + o.fc, d = ket.aob + o.fd;
                    out[((a * n + b) * n + c) * n + d] = v;
                    out[((b * n + a) * n + c) * n + d] = v;
                    out[((a * n + b) * n + d) * n + c] = v;
                    out[((b * n + a) * n + d) * n + c] = v;
                    out[((c * n + d) * n + a) * n + b] = v;
                    out[((d * n + c) * n + a) * n + b] = v;
                    out[((c * n + d) * n + b) * n + a] = v;
                    out[((d * n + c) * n + b) * n + a] = v;
                }
            } else {
                double s = 1.0;
                if (bra.sa == bra.sb) s *= 0.5;
                if (ket.sa == ket.sb) s *= 0.5;
                if (brap == ketp) s *= 0.5;
                for (int e = lane; e < nket_rows; e += 64) mcd[e] = 0.0;
                __syncthreads();
                for (int f = lane; f < S.nf; f += 64) {
                    const HLOut o = hl_out(S, f);
                    const size_t a = bra.aoa + o.fa, b = bra.aob + o.fb, c = ket.aoa + o.fc, d = ket.aob + o.fd;
                    const double v = s * scr[f];
                    atomicAdd(&mab[o.fa * S.nfb + o.fb], 2.0 * v * D[c * n + d]);
                    atomicAdd(&mcd[o.fc * S.nfd + o.fd], 2.0 * v * D[a * n + b]);
                    atomicAdd(&mac[o.fa * S.nfc + o.fc], -0.5 * v * D[b * n + d]);
                    atomicAdd(&mbc[o.fb * S.nfc + o.fc], -0.5 * v * D[a * n + d]);
                    atomicAdd(&mad[o.fa * S.nfd + o.fd], -0.5 * v * D[b * n + c]);
                    atomicAdd(&mbd[o.fb * S.nfd + o.fd], -0.5 * v * D[a * n + c]);
                }
                __syncthreads();
                for (int e = lane; e < nket_rows; e += 64) {
                    size_t row, col;
                    if (e < ncd) {
                        row = ket.aoa + e / S.nfd;
                        col = ket.aob + e % S.nfd;
                    } else if (e < ncd + nac) {
                        const int x = e - ncd;
                        row = bra.aoa + x / S.nfc;
                        col = ket.aoa + x % S.nfc;
                    } else if (e < ncd + nac + nbc) {
                        const int x = e - ncd - nac;
                        row = bra.aob + x / S.nfc;
                        col = ket.aoa + x % S.nfc;
                    } else if (e < ncd + nac + nbc + nad) {
                        const int x = e - ncd - nac - nbc;
                        row = bra.aoa + x / S.nfd;
                        col = ket.aob + x % S.nfd;
                    } else {
                        const int x = e - ncd - nac - nbc - nad;
                        row = bra.aob + x / S.nfd;
                        col = ket.aob + x % S.nfd;
                    }
                    atomicAdd(&M[row * n + col], mcd[e]);
                }
                __syncthreads();
            }
        }
        if constexpr (MODE == 3) {
            for (int e = lane; e < nab; e += 64)
                atomicAdd(&M[(bra.aoa + e / S.nfb) * n + bra.aob + e % S.nfb], mab[e]);
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------- adjoint --
// Adjoint of E2 = 1/2 tr(P G[P]) for the classes with an f or g shell, the
// run-time shaped form of grad_kernel (grad.hip; same energy weights
// W_f = s (4 P_ab P_cd - P_ac P_bd - P_ad P_bc), same derivative identities,
// same task lists: side 0 the Fock build's bra-major tasks, side 1 the
// ket-major ones).  The differentiated pair X (the task's fixed pair) takes
// the bra role of a recurrence raised by 2 on both its centres; rb roots of
// its (root, axis) chains at a time are built in LDS by lanes c < 3 rb, then
// every lane contracts its outputs f = lane + 64 k.  The weights of the
// current quartet live in the workgroup's global scratch slice (up to 50625
// of them).
// Shape of the raised recurrence of class (xa xb | ya yb), X = (xa, xb) raised.
static int hl_groups(const HLShape &S, int ntask, size_t scr_len);

HLShape hl_grad_shape(int xa, int xb, int ya, int yb)
{
    HLShape S;
    S.la = xa + 2;  // HRR b -> a runs to a power xa + 2
    S.lb = xb + 2;  // the d -> c HRR keeps b powers up to xb + 2
    S.lc = ya;
    S.ld = yb;
    S.nfa = ncart(xa);
    S.nfb = ncart(xb);
    S.nfc = ncart(ya);
    S.nfd = ncart(yb);
    S.nf = S.nfa * S.nfb * S.nfc * S.nfd;
    S.nmax = xa + xb + 2;
    S.mmax = ya + yb;
    S.nroots = (xa + xb + ya + yb + 2) / 2 + 1;
    S.dk = xa + 3;
    S.dl = S.dk * (ya + 1);
    S.dj = S.dl * (S.mmax + 1);
    S.gs = S.dj * (S.nmax + 1);
    const int per_root = 3 * S.gs * (int)sizeof(double);
    S.rb = std::max(1, std::min(S.nroots, HL_GRAD_CHAIN_LDS / per_root));
    return S;
}

__global__ __launch_bounds__(64) void hl_grad_kernel(HLShape R, int la, int lb, int lc, int ld, int side,
                                                     const JKTask *__restrict__ tasks, int ntask, DevBasis B,
                                                     const int2 *__restrict__ pp_idx,
                                                     const int2 *__restrict__ shell_info,
                                                     const double *__restrict__ coef, const double *__restrict__ D,
                                                     int nao, double *__restrict__ gatom, double *__restrict__ dprim,
                                                     double *__restrict__ scr_all)
{
    extern __shared__ double g[];
    const int lane = threadIdx.x;
    const int nfa = ncart(la), nfb = ncart(lb), nfc = ncart(lc), nfd = ncart(ld);
    const int nf = nfa * nfb * nfc * nfd;
    double *W = scr_all + (size_t)blockIdx.x * nf;
    const size_t n = (size_t)nao;
    for (int ti = blockIdx.x; ti < ntask; ti += gridDim.x) {
        const JKTask t = tasks[ti];
        const PairView sd = load_pair(B, t.bra);  // the differentiated pair
        const int2 si0 = shell_info[sd.sa], si1 = shell_info[sd.sb];
        int o_begin = 0, o_end = sd.npp;
        if (side == 0) task_prims(t, sd.npp, o_begin, o_end);
        double gx0[3] = {0.0, 0.0, 0.0}, gx1[3] = {0.0, 0.0, 0.0};
        for (int kk = 0; kk < t.nket; kk++) {
            const int otp = t.ket0 + kk;
            const PairView ot = load_pair(B, otp);
            const PairView &br = side == 0 ? sd : ot;  // canonical bra (a b|
            const PairView &kt = side == 0 ? ot : sd;  // canonical ket |c d)
            double s = 1.0;
            if (br.sa == br.sb) s *= 0.5;
            if (kt.sa == kt.sb) s *= 0.5;
            if (t.bra == otp) s *= 0.5;
            for (int f = lane; f < nf; f += 64) {
                const int fa = f % nfa, t1 = f / nfa, fb = t1 % nfb, t2 = t1 / nfb, fc = t2 % nfc, fd = t2 / nfc;
                const size_t a = br.aoa + fa, b = br.aob + fb, c = kt.aoa + fc, d = kt.aob + fd;
                W[f] = s * (4.0 * D[a * n + b] * D[c * n + d] - D[a * n + c] * D[b * n + d] - D[a * n + d] * D[b * n + c]);
            }
            const double rxy[3] = {sd.rb[0] - ot.rb[0], sd.rb[1] - ot.rb[1], sd.rb[2] - ot.rb[2]};
            for (int o = o_begin; o < o_end; o++) {
                const PrimPair po = B.pp[sd.off + o];
                const double al0 = po.pad0, al1 = po.pad1;  // exponents of the two centres of X
                double e0 = 0.0, e1 = 0.0, val = 0.0;
                for (int i = 0; i < ot.npp; i++) {
                    const PrimPair pi = B.pp[ot.off + i];
                    const double aij = po.e, akl = pi.e;
                    const double rsq = rsqrt_pos(aij + akl);
                    const double rs = rsq * rsq;
                    const double a0 = aij * akl * rs;
                    const double pq[3] = {po.pbx - pi.pbx + rxy[0], po.pby - pi.pby + rxy[1], po.pbz - pi.pbz + rxy[2]};
                    const double rr = pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2];
                    const double fac1 = po.K * pi.K * rsq;
                    const double x = a0 * rr;
                    for (int r0 = 0; r0 < R.nroots; r0 += R.rb) {
                        const int nb = min(R.rb, R.nroots - r0);
                        if (lane < 3 * nb) {
                            const int r = r0 + lane / 3, ax = lane - 3 * (lane / 3);
                            double t2, wr;
                            rys_one_rt(R.nroots, x, r, t2, wr);
                            const double tr = t2 * rs;
                            const double b00 = 0.5 * tr;
                            const double b10 = po.h * (1.0 - akl * tr);
                            const double b01 = pi.h * (1.0 - aij * tr);
                            const double pqa = sel3(ax, pq[0], pq[1], pq[2]);
                            const double cn = sel3(ax, po.pbx, po.pby, po.pbz) - akl * tr * pqa;
                            const double cm = sel3(ax, pi.pbx, pi.pby, pi.pbz) + aij * tr * pqa;
                            double *gc = g + lane * R.gs;
                            gc[0] = (ax == 2) ? wr * fac1 : 1.0;
                            vrr_rt(gc, R, cn, cm, b10, b01, b00);
                            hrr_rt(gc, R, sel3(ax, sd.rab[0], sd.rab[1], sd.rab[2]),
                                   sel3(ax, ot.rab[0], ot.rab[1], ot.rab[2]));
                        }
                        __syncthreads();
                        for (int f = lane; f < nf; f += 64) {
                            const double Wk = W[f];
                            const int fa = f % nfa, t1 = f / nfa, fb = t1 % nfb, t2 = t1 / nfb, fc = t2 % nfc,
                                      fd = t2 / nfc;
                            int bx[3], n0[3], n1[3];
#pragma unroll
                            for (int xx = 0; xx < 3; xx++) {
                                const int ia = HCART[la][fa][xx], ib = HCART[lb][fb][xx], ic = HCART[lc][fc][xx],
                                          id = HCART[ld][fd][xx];
                                const int xa = side == 0 ? ia : ic, xb = side == 0 ? ib : id;
                                const int ya = side == 0 ? ic : ia, yb = side == 0 ? id : ib;
                                bx[xx] = xa + ya * R.dk + yb * R.dl + xb * R.dj;
                                n0[xx] = xa;
                                n1[xx] = xb;
                            }
                            for (int c = 0; c < nb; c++) {
                                const double *G0 = g + (3 * c) * R.gs, *G1 = G0 + R.gs, *G2 = G1 + R.gs;
                                const double I[3] = {G0[bx[0]], G1[bx[1]], G2[bx[2]]};
                                val += Wk * I[0] * I[1] * I[2];
#pragma unroll
                                for (int xx = 0; xx < 3; xx++) {
                                    const double *Gx = xx == 0 ? G0 : (xx == 1 ? G1 : G2);
                                    const double o1 = Wk * I[(xx + 1) % 3] * I[(xx + 2) % 3];
                                    double d0 = 2.0 * al0 * Gx[bx[xx] + 1];
                                    if (n0[xx]) d0 -= n0[xx] * Gx[bx[xx] - 1];
                                    double d1 = 2.0 * al1 * Gx[bx[xx] + R.dj];
                                    if (n1[xx]) d1 -= n1[xx] * Gx[bx[xx] - R.dj];
                                    gx0[xx] += d0 * o1;
                                    gx1[xx] += d1 * o1;
                                    e0 -= Gx[bx[xx] + 2] * o1;
                                    e1 -= Gx[bx[xx] + 2 * R.dj] * o1;
                                }
                            }
                        }
                        __syncthreads();
                    }
                }
                e0 = wave_sum64(e0);
                e1 = wave_sum64(e1);
                val = wave_sum64(val);
                if (lane == 0) {
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
        for (int xx = 0; xx < 3; xx++) {
            const double v0 = wave_sum64(gx0[xx]);
            const double v1 = wave_sum64(gx1[xx]);
            if (lane == 0) {
                atomicAdd(&gatom[3 * si0.x + xx], v0);
                atomicAdd(&gatom[3 * si1.x + xx], v1);
            }
        }
    }
}

hipError_t launch_hl_grad(int la, int lb, int lc, int ld, int side, const JKTask *tasks, int ntask, const DevBasis &B,
                          const int2 *pp_idx, const int2 *shell_info, const double *coef, const double *D, int nao,
                          double *gatom, double *dprim, double *scr, size_t scr_len, hipStream_t st)
{
    if (ntask <= 0) return hipSuccess;
    if (side != 0 && side != 1) return hipErrorInvalidValue;
    if (la < 0 || la > lb || lb > HL_LMAX || lc < 0 || lc > ld || ld > HL_LMAX) return hipErrorInvalidValue;
    const HLShape R = side == 0 ? hl_grad_shape(la, lb, lc, ld) : hl_grad_shape(lc, ld, la, lb);
    if (R.nroots > RYS_NMAX || 3 * R.rb > 64) return hipErrorInvalidValue;
    const HLShape O = hl_shape(la, lb, lc, ld);  // outputs: the unraised class
    const int grid = hl_groups(O, ntask, scr_len);
    if (grid <= 0) return hipErrorInvalidValue;
    const size_t lds = sizeof(double) * 3 * (size_t)R.rb * R.gs;
    hipLaunchKernelGGL(hl_grad_kernel, dim3(grid), dim3(64), lds, st, R, la, lb, lc, ld, side, tasks, ntask, B,
                       pp_idx, shell_info, coef, D, nao, gatom, dprim, scr);
    return hipGetLastError();
}

static int hl_groups(const HLShape &S, int ntask, size_t scr_len)
{
    const size_t by_scr = scr_len / std::max<size_t>(1, (size_t)S.nf);
    return (int)std::max<size_t>(0, std::min<size_t>({(size_t)ntask, by_scr, (size_t)HL_MAX_GROUPS}));
}

static size_t hl_lds(const HLShape &S) { return sizeof(double) * 3 * (size_t)S.rb * S.gs; }

static bool hl_valid(const HLShape &S)
{
    return S.la >= 0 && S.la <= S.lb && S.lb <= HL_LMAX && S.lc >= 0 && S.lc <= S.ld && S.ld <= HL_LMAX &&
           S.nroots <= RYS_NMAX && 3 * S.rb <= 64;
}

hipError_t launch_hl_jk(const HLShape &S, const JKTask *tasks, int ntask, const unsigned char *flags, int accept,
                        const DevBasis &B, const double *D, double *M, int nao, double *scr, size_t scr_len,
                        hipStream_t st)
{
    if (ntask <= 0) return hipSuccess;
    if (!hl_valid(S)) return hipErrorInvalidValue;
    const int grid = hl_groups(S, ntask, scr_len);
    if (grid <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((hl_kernel<3>), dim3(grid), dim3(64), hl_lds(S), st, S, tasks, (const ListTask *)nullptr,
                       ntask, flags, accept, B, D, M, (double *)nullptr, nao, scr);
    return hipGetLastError();
}

hipError_t launch_hl_list(const HLShape &S, const ListTask *tasks, int ntask, const DevBasis &B, double *out, int mode,
                          int nao, double *scr, size_t scr_len, hipStream_t st)
{
    if (ntask <= 0) return hipSuccess;
    if (!hl_valid(S)) return hipErrorInvalidValue;
    if (mode == 1 && (S.la != S.lc || S.lb != S.ld)) return hipErrorInvalidValue;
    const int grid = hl_groups(S, ntask, scr_len);
    if (grid <= 0) return hipErrorInvalidValue;
    const size_t lds = hl_lds(S);
#define HL_LAUNCH(m)                                                                                                   \
    hipLaunchKernelGGL((hl_kernel<m>), dim3(grid), dim3(64), lds, st, S, (const JKTask *)nullptr, tasks, ntask,         \
                       (const unsigned char *)nullptr, 0, B, (const double *)nullptr, (double *)nullptr, out, nao, scr)
    if (mode == 0) HL_LAUNCH(0);
    else if (mode == 1) HL_LAUNCH(1);
    else if (mode == 2) HL_LAUNCH(2);
    else return hipErrorInvalidValue;
#undef HL_LAUNCH
    return hipGetLastError();
}

}  // namespace librint
