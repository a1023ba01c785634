// HIP kernels of the two-electron path for gfx950.
//
//   jk_kernel       fused Rys ERI + J/K contraction over Schwarz-screened
//                   canonical shell quartets (the Fock build); one wavefront
//                   per task = one bra shell pair x up to 64 ket shell pairs of
//                   one bin, so every lane runs the same angular class and the
//                   same primitive counts.
//   eri_list_kernel explicit quartet list -> ERI blocks (Schwarz diagonal,
//                   sampled blocks, full tensor scatter).
//   symmetrize_kernel  G = M + M^T.
//
// Reference semantics: src/scf.rs:316-410 (integral2e_fock), src/eri.rs (the
// per-quartet integral).  The 8-fold permutational symmetry and the Schwarz
// screen are additions of this implementation.
#include <hip/hip_runtime.h>

#include "eri_core.h"
#include "kernels.h"

namespace librint {

// ---------------------------------------------------------------- helpers --
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

struct PairView {
    int sa, sb, off, npp;
    int aoa, aob;
    double rb[3], rab[3];
};

__device__ __forceinline__ PairView load_pair(const DevBasis &B, int p)
{
    PairView v;
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

template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void contracted_fp64(double *eri, const DevBasis &B, const PairView &bra, const PairView &ket)
{
    using C = Class<LA, LB, LC, LD>;
#pragma unroll
    for (int f = 0; f < C::NF; f++) eri[f] = 0.0;
    const double rbd[3] = {bra.rb[0] - ket.rb[0], bra.rb[1] - ket.rb[1], bra.rb[2] - ket.rb[2]};
    for (int ib = 0; ib < bra.npp; ib++) {
        const PrimPair p = B.pp[bra.off + ib];
        for (int ik = 0; ik < ket.npp; ik++) {
            const PrimPair q = B.pp[ket.off + ik];
            prim_quartet_fp64<LA, LB, LC, LD>(eri, p, q, rbd, bra.rab, ket.rab);
        }
    }
}

// ------------------------------------------------------------ J/K kernel --
// M accumulates A - K_half so that G = M + M^T (see symmetrize_kernel):
//   for one canonical quartet with integral v and shell-degeneracy scale s
//     M[a][b] += 2 s D[c][d] v      M[c][d] += 2 s D[a][b] v
//     M[a][c] -= s/2 D[b][d] v      M[b][c] -= s/2 D[a][d] v
//     M[a][d] -= s/2 D[b][c] v      M[b][d] -= s/2 D[a][c] v
// which sums the J and K contributions of all 8 index images.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void jk_kernel(const JKTask *__restrict__ tasks, DevBasis B,
                                                const double *__restrict__ D, double *__restrict__ M, int nao)
{
    using C = Class<LA, LB, LC, LD>;
    const JKTask t = tasks[blockIdx.x];
    const int lane = threadIdx.x;
    const bool active = lane < t.nket;
    const int ketp = t.ket0 + (active ? lane : 0);
    const PairView bra = load_pair(B, t.bra);
    const PairView ket = load_pair(B, ketp);

    double eri[C::NF];
    contracted_fp64<LA, LB, LC, LD>(eri, B, bra, ket);

    double s = active ? 1.0 : 0.0;
    if (bra.sa == bra.sb) s *= 0.5;
    if (ket.sa == ket.sb) s *= 0.5;
    if (t.bra == ketp) s *= 0.5;

    const size_t n = (size_t)nao;
    // bra-bra block: reduced over the wavefront, one atomic per element
#pragma unroll
    for (int fa = 0; fa < C::NFA; fa++)
#pragma unroll
        for (int fb = 0; fb < C::NFB; fb++) {
            double acc = 0.0;
#pragma unroll
            for (int fd = 0; fd < C::NFD; fd++)
#pragma unroll
                for (int fc = 0; fc < C::NFC; fc++)
                    acc += D[(ket.aoa + fc) * n + ket.aob + fd] *
                           eri[fa + C::NFA * (fb + C::NFB * (fc + C::NFC * fd))];
            acc = wave_sum(2.0 * s * acc);
            if (lane == 0) atomicAdd(&M[(bra.aoa + fa) * n + bra.aob + fb], acc);
        }
    if (!active) return;
    // ket-ket block
#pragma unroll
    for (int fc = 0; fc < C::NFC; fc++)
#pragma unroll
        for (int fd = 0; fd < C::NFD; fd++) {
            double acc = 0.0;
#pragma unroll
            for (int fb = 0; fb < C::NFB; fb++)
#pragma unroll
                for (int fa = 0; fa < C::NFA; fa++)
                    acc += D[(bra.aoa + fa) * n + bra.aob + fb] *
                           eri[fa + C::NFA * (fb + C::NFB * (fc + C::NFC * fd))];
            atomicAdd(&M[(ket.aoa + fc) * n + ket.aob + fd], 2.0 * s * acc);
        }
    const double h = -0.5 * s;
    // exchange: (a,c) and (b,c)
#pragma unroll
    for (int fc = 0; fc < C::NFC; fc++) {
#pragma unroll
        for (int fa = 0; fa < C::NFA; fa++) {
            double acc = 0.0;
#pragma unroll
            for (int fd = 0; fd < C::NFD; fd++)
#pragma unroll
                for (int fb = 0; fb < C::NFB; fb++)
                    acc += D[(bra.aob + fb) * n + ket.aob + fd] *
                           eri[fa + C::NFA * (fb + C::NFB * (fc + C::NFC * fd))];
            atomicAdd(&M[(bra.aoa + fa) * n + ket.aoa + fc], h * acc);
        }
#pragma unroll
        for (int fb = 0; fb < C::NFB; fb++) {
            double acc = 0.0;
#pragma unroll
            for (int fd = 0; fd < C::NFD; fd++)
#pragma unroll
                for (int fa = 0; fa < C::NFA; fa++)
                    acc += D[(bra.aoa + fa) * n + ket.aob + fd] *
                           eri[fa + C::NFA * (fb + C::NFB * (fc + C::NFC * fd))];
            atomicAdd(&M[(bra.aob + fb) * n + ket.aoa + fc], h * acc);
        }
    }
    // exchange: (a,d) and (b,d)
#pragma unroll
    for (int fd = 0; fd < C::NFD; fd++) {
#pragma unroll
        for (int fa = 0; fa < C::NFA; fa++) {
            double acc = 0.0;
#pragma unroll
            for (int fc = 0; fc < C::NFC; fc++)
#pragma unroll
                for (int fb = 0; fb < C::NFB; fb++)
                    acc += D[(bra.aob + fb) * n + ket.aoa + fc] *
                           eri[fa + C::NFA * (fb + C::NFB * (fc + C::NFC * fd))];
            atomicAdd(&M[(bra.aoa + fa) * n + ket.aob + fd], h * acc);
        }
#pragma unroll
        for (int fb = 0; fb < C::NFB; fb++) {
            double acc = 0.0;
#pragma unroll
            for (int fc = 0; fc < C::NFC; fc++)
#pragma unroll
                for (int fa = 0; fa < C::NFA; fa++)
                    acc += D[(bra.aoa + fa) * n + ket.aoa + fc] *
                           eri[fa + C::NFA * (fb + C::NFB * (fc + C::NFC * fd))];
            atomicAdd(&M[(bra.aob + fb) * n + ket.aob + fd], h * acc);
        }
    }
}

// ---------------------------------------------------------- list kernel --
// mode 0: write the block to out[task.out .. + NF) (a fastest)
// mode 1: write sqrt(max_ab |(ab|ab)|) to out[task.out] (Schwarz diagonal)
// mode 2: scatter the block into the nao^4 tensor out at all 8 images
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void eri_list_kernel(const ListTask *__restrict__ tasks, int ntask, DevBasis B,
                                                      double *__restrict__ out, int mode, int nao)
{
    using C = Class<LA, LB, LC, LD>;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= ntask) return;
    const ListTask t = tasks[tid];
    const PairView bra = load_pair(B, t.bra);
    const PairView ket = load_pair(B, t.ket);
    double eri[C::NF];
    contracted_fp64<LA, LB, LC, LD>(eri, B, bra, ket);
    if (mode == 0) {
#pragma unroll
        for (int f = 0; f < C::NF; f++) out[t.out + f] = eri[f];
    } else if (mode == 1) {
        double m = 0.0;
        if constexpr (LA == LC && LB == LD) {
#pragma unroll
            for (int fb = 0; fb < C::NFB; fb++)
#pragma unroll
                for (int fa = 0; fa < C::NFA; fa++) m = fmax(m, fabs(eri[fa + C::NFA * (fb + C::NFB * (fa + C::NFC * fb))]));
        }
        out[t.out] = sqrt(m);
    } else {
        const size_t n = (size_t)nao;
#pragma unroll
        for (int fd = 0; fd < C::NFD; fd++)
#pragma unroll
            for (int fc = 0; fc < C::NFC; fc++)
#pragma unroll
                for (int fb = 0; fb < C::NFB; fb++)
#pragma unroll
                    for (int fa = 0; fa < C::NFA; fa++) {
                        const double v = eri[fa + C::NFA * (fb + C::NFB * (fc + C::NFC * fd))];
                        const size_t a = bra.aoa + fa, b = bra.aob + fb, c = ket.aoa + fc, d = ket.aob + fd;
                        out[((a * n + b) * n + c) * n + d] = v;
                        out[((b * n + a) * n + c) * n + d] = v;
                        out[((a * n + b) * n + d) * n + c] = v;
                        out[((b * n + a) * n + d) * n + c] = v;
                        out[((c * n + d) * n + a) * n + b] = v;
                        out[((d * n + c) * n + a) * n + b] = v;
                        out[((c * n + d) * n + b) * n + a] = v;
                        out[((d * n + c) * n + b) * n + a] = v;
                    }
    }
}

__global__ void symmetrize_kernel(const double *__restrict__ M, double *__restrict__ G, int nao)
{
    const size_t n = (size_t)nao;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const size_t i = idx / n, j = idx % n;
    G[idx] = M[i * n + j] + M[j * n + i];
}

// ---------------------------------------------------------- dispatch --
// Pair classes (la <= lb): 0:(0,0) 1:(0,1) 2:(1,1) 3:(0,2) 4:(1,2) 5:(2,2)
#define LR_FOR_CLASS_PAIRS(X)                                                                                          \
    X(0, 0, 0, 0) X(0, 1, 0, 0) X(0, 1, 0, 1) X(1, 1, 0, 0) X(1, 1, 0, 1) X(1, 1, 1, 1)

static int cls_index(int la, int lb)
{
    static const int tab[3][3] = {{0, 1, 3}, {-1, 2, 4}, {-1, -1, 5}};
    return tab[la][lb];
}

bool class_supported(int la, int lb, int lc, int ld)
{
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d) return true;
    LR_FOR_CLASS_PAIRS(X)
#undef X
    return false;
}

hipError_t launch_jk(int la, int lb, int lc, int ld, const JKTask *tasks, int ntask, const DevBasis &B,
                     const double *D, double *M, int nao, hipStream_t st)
{
    if (ntask == 0) return hipSuccess;
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d) {                                                                    \
        hipLaunchKernelGGL((jk_kernel<a, b, c, d>), dim3(ntask), dim3(64), 0, st, tasks, B, D, M, nao);               \
        return hipGetLastError();                                                                                      \
    }
    LR_FOR_CLASS_PAIRS(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_list(int la, int lb, int lc, int ld, const ListTask *tasks, int ntask, const DevBasis &B,
                       double *out, int mode, int nao, hipStream_t st)
{
    if (ntask == 0) return hipSuccess;
    const int nblk = (ntask + 63) / 64;
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d) {                                                                    \
        hipLaunchKernelGGL((eri_list_kernel<a, b, c, d>), dim3(nblk), dim3(64), 0, st, tasks, ntask, B, out, mode,    \
                           nao);                                                                                       \
        return hipGetLastError();                                                                                      \
    }
    LR_FOR_CLASS_PAIRS(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_symmetrize(const double *M, double *G, int nao, hipStream_t st)
{
    const size_t nn = (size_t)nao * nao;
    hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, M, G, nao);
    return hipGetLastError();
}

} // namespace librint
