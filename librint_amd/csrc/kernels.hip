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

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

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
//
// Task flags (per build, from classify_kernel): 0 FP64, 1 FP32 recurrences,
// 2 skipped (density-weighted bound below the screening threshold).  A
// launch processes the tasks whose flag bit is set in `accept`; flags ==
// nullptr accepts every task.
//
// Working precision T: the primitive loop, Rys evaluation and the
// contracted block run in T; the J/K contraction and every accumulation
// into M are FP64.
template <int LA, int LB, int LC, int LD, typename T, int SD = -1, int S0 = 0>
__device__ __forceinline__ void contract_block(const JKTask &t, const PairView &bra, const PairView &ket, int ketp,
                                               const DevBasis &B, T *eri, const T *rt)
{
    using C = Class<LA, LB, LC, LD>;
#pragma unroll
    for (int f = 0; f < Chunk<LA, LB, LC, LD, SD, S0>::NFL; f++) eri[f] = (T)0;
    const T rbd[3] = {(T)(bra.rb[0] - ket.rb[0]), (T)(bra.rb[1] - ket.rb[1]), (T)(bra.rb[2] - ket.rb[2])};
    const T rab[3] = {(T)bra.rab[0], (T)bra.rab[1], (T)bra.rab[2]};
    const T rcd[3] = {(T)ket.rab[0], (T)ket.rab[1], (T)ket.rab[2]};
    // every ket of a task comes from one bin, so the primitive-pair count is
    // the same in all lanes: make the loop bounds scalar for the compiler
    int b0, b1;
    task_prims(t, bra.npp, b0, b1);
    const int knpp = __builtin_amdgcn_readfirstlane(ket.npp);
    const int bnpp = __builtin_amdgcn_readfirstlane(b1 - b0);
    const int boff = __builtin_amdgcn_readfirstlane(bra.off + b0);
    // Light classes (s on both HRR targets, NF <= 9) are software pipelined:
    // the next primitive-pair records load while the current primitive
    // quartet is evaluated, hiding their latency.  Heavier classes keep the
    // plain loop: the extra registers would cost them occupancy.
    constexpr bool PIPE = LA == 0 && LC == 0 && C::NF <= 9;
    // (01|00) keeps its bra loop rolled: with the occupancy request on its
    // kernel (jk_group_kernel) it then fits 128 VGPRs with 5 spilled
    // (4 wavefronts per SIMD instead of 3; measured 0.500 -> 0.471 ms), where
    // the unrolled loop spills 20.  The other light classes gain from the
    // unrolled pair.
    constexpr int BUNROLL = (LA == 0 && LB == 1 && LC == 0 && LD == 0) ? 1 : 2;
    if constexpr (PIPE) {
        PrimPair qn = B.pp0[ketp];
        for (int ik = 0; ik < knpp; ik++) {
            const PrimPair q = qn;
            if (ik + 1 < knpp) qn = B.pp[ket.off + ik + 1];
            const T qd[3] = {(T)q.pbx - rbd[0], (T)q.pby - rbd[1], (T)q.pbz - rbd[2]};
            PrimPair pn = B.pp[boff];
#pragma unroll BUNROLL
            for (int ib = 0; ib < bnpp; ib++) {
                const PrimPair p = pn;
                if (ib + 1 < bnpp) pn = B.pp[boff + ib + 1];
                prim_quartet_lds<LA, LB, LC, LD, T, SD, S0>(eri, p, q, qd, rab, rcd, rt);
            }
        }
    } else {
        for (int ik = 0; ik < knpp; ik++) {
            const PrimPair q = ik == 0 ? B.pp0[ketp] : B.pp[ket.off + ik];
            const T qd[3] = {(T)q.pbx - rbd[0], (T)q.pby - rbd[1], (T)q.pbz - rbd[2]};
            for (int ib = 0; ib < bnpp; ib++) {
                const PrimPair p = B.pp[boff + ib];
                prim_quartet_lds<LA, LB, LC, LD, T, SD, S0>(eri, p, q, qd, rab, rcd, rt);
            }
        }
    }
}

template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ void jk_task(const JKTask &t, const DevBasis &B, const double *__restrict__ D,
                                        double *__restrict__ M, int nao, int lane, const T *rt, double *jslab)
{
    using C = Class<LA, LB, LC, LD>;
    const bool active = lane < t.nket;
    const int ketp = t.ket0 + (active ? lane : 0);
    const PairView bra = load_pair(B, t.bra);
    const PairView ket = load_pair(B, ketp);
    T eri[C::NF];
    contract_block<LA, LB, LC, LD, T>(t, bra, ket, ketp, B, eri, rt);

    double s = active ? 1.0 : 0.0;
    if (bra.sa == bra.sb) s *= 0.5;
    if (ket.sa == ket.sb) s *= 0.5;
    if (t.bra == ketp) s *= 0.5;
    const size_t n = (size_t)nao;
#define ERI_(fa, fb, fc, fd) ((double)eri[(fa) + C::NFA * ((fb) + C::NFB * ((fc) + C::NFC * (fd)))])
    // bra-bra block: reduced over the wavefront, one atomic per element
#pragma unroll
    for (int fa = 0; fa < C::NFA; fa++)
#pragma unroll
        for (int fb = 0; fb < C::NFB; fb++) {
            double acc = 0.0;
#pragma unroll
            for (int fd = 0; fd < C::NFD; fd++)
#pragma unroll
                for (int fc = 0; fc < C::NFC; fc++) acc += D[(ket.aoa + fc) * n + ket.aob + fd] * ERI_(fa, fb, fc, fd);
            acc = wave_sum(2.0 * s * acc);
            if (lane == 0) atomicAdd(&M[(bra.aoa + fa) * n + bra.aob + fb], acc);
        }
    if (!active) return;
#pragma unroll
    for (int fc = 0; fc < C::NFC; fc++)
#pragma unroll
        for (int fd = 0; fd < C::NFD; fd++) {
            double acc = 0.0;
#pragma unroll
            for (int fb = 0; fb < C::NFB; fb++)
#pragma unroll
                for (int fa = 0; fa < C::NFA; fa++) acc += D[(bra.aoa + fa) * n + bra.aob + fb] * ERI_(fa, fb, fc, fd);
            if (jslab)
                jslab[(fc * C::NFD + fd) * 64 + lane] = 2.0 * s * acc;
            else
                atomicAdd(&M[(ket.aoa + fc) * n + ket.aob + fd], 2.0 * s * acc);
        }
    const double h = -0.5 * s;
#pragma unroll
    for (int fc = 0; fc < C::NFC; fc++) {
#pragma unroll
        for (int fa = 0; fa < C::NFA; fa++) {
            double acc = 0.0;
#pragma unroll
            for (int fd = 0; fd < C::NFD; fd++)
#pragma unroll
                for (int fb = 0; fb < C::NFB; fb++) acc += D[(bra.aob + fb) * n + ket.aob + fd] * ERI_(fa, fb, fc, fd);
            atomicAdd(&M[(bra.aoa + fa) * n + ket.aoa + fc], h * acc);
        }
#pragma unroll
        for (int fb = 0; fb < C::NFB; fb++) {
            double acc = 0.0;
#pragma unroll
            for (int fd = 0; fd < C::NFD; fd++)
#pragma unroll
                for (int fa = 0; fa < C::NFA; fa++) acc += D[(bra.aoa + fa) * n + ket.aob + fd] * ERI_(fa, fb, fc, fd);
            atomicAdd(&M[(bra.aob + fb) * n + ket.aoa + fc], h * acc);
        }
    }
#pragma unroll
    for (int fd = 0; fd < C::NFD; fd++) {
#pragma unroll
        for (int fa = 0; fa < C::NFA; fa++) {
            double acc = 0.0;
#pragma unroll
            for (int fc = 0; fc < C::NFC; fc++)
#pragma unroll
                for (int fb = 0; fb < C::NFB; fb++) acc += D[(bra.aob + fb) * n + ket.aoa + fc] * ERI_(fa, fb, fc, fd);
            atomicAdd(&M[(bra.aoa + fa) * n + ket.aob + fd], h * acc);
        }
#pragma unroll
        for (int fb = 0; fb < C::NFB; fb++) {
            double acc = 0.0;
#pragma unroll
            for (int fc = 0; fc < C::NFC; fc++)
#pragma unroll
                for (int fa = 0; fa < C::NFA; fa++) acc += D[(bra.aoa + fa) * n + ket.aoa + fc] * ERI_(fa, fb, fc, fd);
            atomicAdd(&M[(bra.aob + fb) * n + ket.aob + fd], h * acc);
        }
    }
#undef ERI_
}

// Persistent form: 256-thread workgroups stage the Rys table block of the
// class's root count in LDS once, then each wavefront pulls tasks from a
// per-launch counter.
template <int LA, int LB, int LC, int LD, typename T>
__global__ __launch_bounds__(256) void jk_kernel(const JKTask *__restrict__ tasks, int ntask, int *__restrict__ counter,
                                                 const unsigned char *__restrict__ flags, int accept, DevBasis B,
                                                 const double *__restrict__ D, double *__restrict__ M, int nao,
                                                 double *__restrict__ jslab, int tfirst)
{
    using C = Class<LA, LB, LC, LD>;
    using R = RysLds<LA, LB, LC, LD, T>;
    __shared__ T rt[R::LEN];
    rys_stage<T, R::N, R::WONLY>(rt);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter, 1);
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= ntask || t < 0) break;
        if (flags && !((accept >> flags[t]) & 1)) continue;
        double *js = jslab ? jslab + (size_t)(t - tfirst) * (64 * C::NFC * C::NFD) : nullptr;
        jk_task<LA, LB, LC, LD, T>(tasks[t], B, D, M, nao, lane, rt, js);
    }
}

// ------------------------------------------------------ grouped J/K kernel --
// One workgroup owns one bra shell pair (a,b) and all its kets of this class
// pair (a "group" of tasks); its 4 wavefronts take the group's tasks in turn.
// Every exchange contribution of the group lands in the bra's own AO rows
// (M[a][*] and M[b][*]), so they are accumulated in LDS (ds_add_f64) and
// flushed once per group with coalesced global atomics; the bra-bra Coulomb
// block is reduced over the whole group.  Only the ket-ket Coulomb block
// (M[c][d], one atomic per element per quartet) goes straight to memory.
// Classes whose contracted block is evaluated in chunks (DIM: the centre
// whose component index selects a chunk, 0 a .. 3 d; N chunks): their block
// (162-324 integrals) does not fit a lane's registers.  Each chunk repeats
// the primitive loop (Rys nodes, VRR, HRR) and keeps only its own outputs,
// so a lane still evaluates a whole quartet.  On the uncontracted lists
// these classes have 1-2 primitive quartets per quartet, so the repeated
// recurrences cost less than giving a quartet a whole wavefront.
template <int LA, int LB, int LC, int LD>
struct Split {
    static constexpr int DIM = -1, N = 1;
};
template <> struct Split<1, 2, 1, 1> { static constexpr int DIM = 1, N = 6; };
template <> struct Split<1, 2, 1, 2> { static constexpr int DIM = 1, N = 6; };
template <> struct Split<2, 2, 1, 1> { static constexpr int DIM = 0, N = 6; };
template <> struct Split<2, 2, 0, 2> { static constexpr int DIM = 0, N = 6; };

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// One task of a chunked class (lane = ket of the task): per chunk, the
// chunk's integrals, then its share of every J/K term.  The bra Coulomb
// entries of a chunk are reduced over the wavefront and added to the
// group's block in LDS (mab_l); the ket Coulomb block is summed over the
// chunks in registers; the exchange terms go to the LDS rows as in the
// whole-block form.  Inactive lanes (past the task's kets) evaluate a valid
// ket with weight 0 and take part in the reductions only.
template <int LA, int LB, int LC, int LD, typename T, bool DROW>
__device__ __forceinline__ void jk_task_chunks(const JKTask &t, const PairView &bra, const PairView &ket, int ketp,
                                               bool active, const DevBasis &B, const double *__restrict__ D,
                                               double *__restrict__ M, int nao, int lane, const T *rt, double *Krow,
                                               double *jslab, const double *Drow, double *mab_l)
{
    using C = Class<LA, LB, LC, LD>;
    using SP = Split<LA, LB, LC, LD>;
    double s = active ? 1.0 : 0.0;
    if (bra.sa == bra.sb) s *= 0.5;
    if (ket.sa == ket.sb) s *= 0.5;
    if (t.bra == ketp) s *= 0.5;
    const double h = -0.5 * s;
    const size_t n = (size_t)nao;
    double *Ka = Krow, *Kb = Krow + C::NFA * n;
    double jcd[C::NFC * C::NFD];
#pragma unroll
    for (int e = 0; e < C::NFC * C::NFD; e++) jcd[e] = 0.0;
    // a run-time loop over the chunks (one compiled body per chunk behind a
    // wavefront-uniform branch): unrolled, the compiler overlaps the chunks
    // and keeps several chunks' recurrence values live
#pragma unroll 1
    for (int cidx = 0; cidx < SP::N; cidx++)
    static_for<0, SP::N>([&](auto ch) {
        constexpr int S0 = decltype(ch)::value;
        if (cidx != S0) return;
        using K = Chunk<LA, LB, LC, LD, SP::DIM, S0>;
        // The density pointers pass through an empty asm statement per
        // chunk, so that each chunk loads its density values again: reusing
        // the first chunk's loads would keep every chunk's values live at
        // once (81-162 doubles), far beyond the registers a lane has.
        const double *Dk = D, *Dr = Drow;
        asm volatile("" : "+s"(Dk));
        if constexpr (DROW) asm volatile("" : "+s"(Dr));
        auto Da = [&](int fa, size_t col) { return DROW ? Dr[fa * n + col] : Dk[(bra.aoa + fa) * n + col]; };
        auto Db = [&](int fb, size_t col) {
            return DROW ? Dr[(C::NFA + fb) * n + col] : Dk[(bra.aob + fb) * n + col];
        };
        T eri[K::NFL];
        contract_block<LA, LB, LC, LD, T, SP::DIM, S0>(t, bra, ket, ketp, B, eri, rt);
#define ERI_(fa, fb, fc, fd) ((double)eri[K::idx(fa, fb, fc, fd)])
#pragma unroll
        for (int fa = K::A0; fa < K::A0 + K::NA; fa++)
#pragma unroll
            for (int fb = K::B0; fb < K::B0 + K::NB; fb++) {
                double acc = 0.0;
#pragma unroll
                for (int fd = K::D0; fd < K::D0 + K::ND; fd++)
#pragma unroll
                    for (int fc = K::C0; fc < K::C0 + K::NC; fc++)
                        acc += Dk[(ket.aoa + fc) * n + ket.aob + fd] * ERI_(fa, fb, fc, fd);
                const double v = wave_sum(2.0 * s * acc);
                if (lane == 0) atomicAdd(&mab_l[fa * C::NFB + fb], v);
            }
#pragma unroll
        for (int fc = K::C0; fc < K::C0 + K::NC; fc++)
#pragma unroll
            for (int fd = K::D0; fd < K::D0 + K::ND; fd++) {
                double acc = 0.0;
#pragma unroll
                for (int fb = K::B0; fb < K::B0 + K::NB; fb++)
#pragma unroll
                    for (int fa = K::A0; fa < K::A0 + K::NA; fa++)
                        acc += Dk[(bra.aoa + fa) * n + bra.aob + fb] * ERI_(fa, fb, fc, fd);
                jcd[fc * C::NFD + fd] += 2.0 * s * acc;
            }
        if (active) {
#pragma unroll
            for (int fc = K::C0; fc < K::C0 + K::NC; fc++) {
#pragma unroll
                for (int fa = K::A0; fa < K::A0 + K::NA; fa++) {
                    double acc = 0.0;
#pragma unroll
                    for (int fd = K::D0; fd < K::D0 + K::ND; fd++)
#pragma unroll
                        for (int fb = K::B0; fb < K::B0 + K::NB; fb++) acc += Db(fb, ket.aob + fd) * ERI_(fa, fb, fc, fd);
                    atomicAdd(&Ka[fa * n + ket.aoa + fc], h * acc);
                }
#pragma unroll
                for (int fb = K::B0; fb < K::B0 + K::NB; fb++) {
                    double acc = 0.0;
#pragma unroll
                    for (int fd = K::D0; fd < K::D0 + K::ND; fd++)
#pragma unroll
                        for (int fa = K::A0; fa < K::A0 + K::NA; fa++) acc += Da(fa, ket.aob + fd) * ERI_(fa, fb, fc, fd);
                    atomicAdd(&Kb[fb * n + ket.aoa + fc], h * acc);
                }
            }
#pragma unroll
            for (int fd = K::D0; fd < K::D0 + K::ND; fd++) {
#pragma unroll
                for (int fa = K::A0; fa < K::A0 + K::NA; fa++) {
                    double acc = 0.0;
#pragma unroll
                    for (int fc = K::C0; fc < K::C0 + K::NC; fc++)
#pragma unroll
                        for (int fb = K::B0; fb < K::B0 + K::NB; fb++) acc += Db(fb, ket.aoa + fc) * ERI_(fa, fb, fc, fd);
                    atomicAdd(&Ka[fa * n + ket.aob + fd], h * acc);
                }
#pragma unroll
                for (int fb = K::B0; fb < K::B0 + K::NB; fb++) {
                    double acc = 0.0;
#pragma unroll
                    for (int fc = K::C0; fc < K::C0 + K::NC; fc++)
#pragma unroll
                        for (int fa = K::A0; fa < K::A0 + K::NA; fa++) acc += Da(fa, ket.aoa + fc) * ERI_(fa, fb, fc, fd);
                    atomicAdd(&Kb[fb * n + ket.aob + fd], h * acc);
                }
            }
        }
#undef ERI_
    });
    if (!active) return;
#pragma unroll
    for (int fc = 0; fc < C::NFC; fc++)
#pragma unroll
        for (int fd = 0; fd < C::NFD; fd++) {
            if (jslab)
                jslab[(fc * C::NFD + fd) * 64 + lane] = jcd[fc * C::NFD + fd];
            else
                atomicAdd(&M[(ket.aoa + fc) * n + ket.aob + fd], jcd[fc * C::NFD + fd]);
        }
}

template <int LA, int LB, int LC, int LD, typename T, bool DROW>
__device__ __forceinline__ void jk_task_rows(const JKTask &t, const PairView &bra, const DevBasis &B,
                                             const double *__restrict__ D, double *__restrict__ M, int nao, int lane,
                                             const T *rt, double *Krow, double *mab, double *jslab,
                                             const double *Drow, double *mab_l)
{
    using C = Class<LA, LB, LC, LD>;
    using SP = Split<LA, LB, LC, LD>;
    const bool active = lane < t.nket;
    const int ketp = t.ket0 + (active ? lane : 0);
    const PairView ket = load_pair(B, ketp);
    if constexpr (SP::N > 1) {
        jk_task_chunks<LA, LB, LC, LD, T, DROW>(t, bra, ket, ketp, active, B, D, M, nao, lane, rt, Krow, jslab, Drow,
                                                mab_l);
        return;
    }
    T eri[C::NF];
    contract_block<LA, LB, LC, LD, T>(t, bra, ket, ketp, B, eri, rt);
    if (!active) return;
    double s = 1.0;
    if (bra.sa == bra.sb) s *= 0.5;
    if (ket.sa == ket.sb) s *= 0.5;
    if (t.bra == ketp) s *= 0.5;
    const size_t n = (size_t)nao;
#define ERI_(fa, fb, fc, fd) ((double)eri[(fa) + C::NFA * ((fb) + C::NFB * ((fc) + C::NFC * (fd)))])
#pragma unroll
    for (int fa = 0; fa < C::NFA; fa++)
#pragma unroll
        for (int fb = 0; fb < C::NFB; fb++) {
            double acc = 0.0;
#pragma unroll
            for (int fd = 0; fd < C::NFD; fd++)
#pragma unroll
                for (int fc = 0; fc < C::NFC; fc++) acc += D[(ket.aoa + fc) * n + ket.aob + fd] * ERI_(fa, fb, fc, fd);
            mab[fa * C::NFB + fb] += 2.0 * s * acc;
        }
#pragma unroll
    for (int fc = 0; fc < C::NFC; fc++)
#pragma unroll
        for (int fd = 0; fd < C::NFD; fd++) {
            double acc = 0.0;
#pragma unroll
            for (int fb = 0; fb < C::NFB; fb++)
#pragma unroll
                for (int fa = 0; fa < C::NFA; fa++) acc += D[(bra.aoa + fa) * n + bra.aob + fb] * ERI_(fa, fb, fc, fd);
            if (jslab)
                jslab[(fc * C::NFD + fd) * 64 + lane] = 2.0 * s * acc;
            else
                atomicAdd(&M[(ket.aoa + fc) * n + ket.aob + fd], 2.0 * s * acc);
        }
    const double h = -0.5 * s;
    double *Ka = Krow, *Kb = Krow + C::NFA * n;
    // density rows a and b: from LDS (DROW) or from D
    auto Da = [&](int fa, size_t col) { return DROW ? Drow[fa * n + col] : D[(bra.aoa + fa) * n + col]; };
    auto Db = [&](int fb, size_t col) {
        return DROW ? Drow[(C::NFA + fb) * n + col] : D[(bra.aob + fb) * n + col];
    };
#pragma unroll
    for (int fc = 0; fc < C::NFC; fc++) {
#pragma unroll
        for (int fa = 0; fa < C::NFA; fa++) {
            double acc = 0.0;
#pragma unroll
            for (int fd = 0; fd < C::NFD; fd++)
#pragma unroll
                for (int fb = 0; fb < C::NFB; fb++) acc += Db(fb, ket.aob + fd) * ERI_(fa, fb, fc, fd);
            atomicAdd(&Ka[fa * n + ket.aoa + fc], h * acc);
        }
#pragma unroll
        for (int fb = 0; fb < C::NFB; fb++) {
            double acc = 0.0;
#pragma unroll
            for (int fd = 0; fd < C::NFD; fd++)
#pragma unroll
                for (int fa = 0; fa < C::NFA; fa++) acc += Da(fa, ket.aob + fd) * ERI_(fa, fb, fc, fd);
            atomicAdd(&Kb[fb * n + ket.aoa + fc], h * acc);
        }
    }
#pragma unroll
    for (int fd = 0; fd < C::NFD; fd++) {
#pragma unroll
        for (int fa = 0; fa < C::NFA; fa++) {
            double acc = 0.0;
#pragma unroll
            for (int fc = 0; fc < C::NFC; fc++)
#pragma unroll
                for (int fb = 0; fb < C::NFB; fb++) acc += Db(fb, ket.aoa + fc) * ERI_(fa, fb, fc, fd);
            atomicAdd(&Ka[fa * n + ket.aob + fd], h * acc);
        }
#pragma unroll
        for (int fb = 0; fb < C::NFB; fb++) {
            double acc = 0.0;
#pragma unroll
            for (int fc = 0; fc < C::NFC; fc++)
#pragma unroll
                for (int fa = 0; fa < C::NFA; fa++) acc += Da(fa, ket.aoa + fc) * ERI_(fa, fb, fc, fd);
            atomicAdd(&Kb[fb * n + ket.aob + fd], h * acc);
        }
    }
#undef ERI_
}

// LDS footprint of the Rys table block in doubles, rounded up to 16 bytes
template <int LA, int LB, int LC, int LD, typename T>
__host__ __device__ constexpr int RysLdsDoubles()
{
    return (int)((RysLds<LA, LB, LC, LD, T>::LEN * sizeof(T) + 15) / 16 * 2);
}

// LDS doubles of the Rys tables a group kernel stages: the FP64 block, and
// with MIXED the FP32 block behind it
template <int LA, int LB, int LC, int LD, bool MIXED>
__host__ __device__ constexpr int GroupTabDoubles()
{
    return RysLdsDoubles<LA, LB, LC, LD, double>() + (MIXED ? RysLdsDoubles<LA, LB, LC, LD, float>() : 0);
}

// gmask: per group, bit f set when the group holds a task with flag f.
// DROW: the bra's density rows (NFA + NFB rows of nao) are staged in LDS
// beside the exchange rows, so the exchange terms read D from LDS.
// MIXED: each task runs FP32 recurrences when its bound
// Q_b max(Dpair[b] Q[ket0], Wqd[ket0]) (MixArgs, wavefront-uniform scalar
// loads) is below tau_mix, else FP64, and records the choice in prec[t]; the
// FP32 path has its own Rys table block in LDS.  Each wavefront loads the
// classification flags of its tasks of a group at once (lane j: its j-th
// task), so the per-task skip test does not wait on memory.
// Minimum wavefronts per SIMD asked of the compiler per class: 4 for (01|00)
// (153 VGPRs would allow 3; 128 with 5 spilled), 5 for (00|00) (102 -> 96
// VGPRs, no spill); no request elsewhere (59 spills for 3 in (01|01)) nor
// with the density rows in LDS (the LDS then allows two workgroups per CU).
// Wavefronts per workgroup.  Without density rows in LDS: 4.  With them:
// 12 for (01|00) (its 73 KB of LDS at nao 992 leaves 2 workgroups of 4 per
// CU, 2 waves per SIMD, while its registers allow 3: one workgroup of 12
// holds 3); 8 for the classes whose registers allow 2 waves per SIMD
// ((01|01), (11|00), (11|01), (02|00), (02|01), (12|00): one workgroup of 8
// holds 2 even where the rows leave room for one workgroup only); 4 else.
template <int LA, int LB, int LC, int LD, bool DROW>
struct GroupNW {
    static constexpr bool TWO = (LA == 0 && LB == 1 && LC == 0 && LD == 1) || (LA == 1 && LB == 1 && LC == 0 && LD == 0) ||
                                (LA == 1 && LB == 1 && LC == 0 && LD == 1) || (LA == 0 && LB == 2 && LC == 0 && LD == 0) ||
                                (LA == 0 && LB == 2 && LC == 0 && LD == 1) || (LA == 1 && LB == 2 && LC == 0 && LD == 0);
    static constexpr int value = !DROW ? 4 : (LA == 0 && LB == 1 && LC == 0 && LD == 0) ? 12 : (TWO ? 8 : 4);
};

template <int LA, int LB, int LC, int LD, bool DROW>
struct GroupWaves {
    static constexpr int value = DROW ? 1
                                 : (LA == 0 && LB == 1 && LC == 0 && LD == 0)   ? 4
                                 : (LA == 0 && LB == 0 && LC == 0 && LD == 0)   ? 5
                                                                                : 1;
};

template <int LA, int LB, int LC, int LD, bool MIXED, bool DROW>
__global__ __launch_bounds__((64 * GroupNW<LA, LB, LC, LD, DROW>::value))
__attribute__((amdgpu_waves_per_eu(GroupWaves<LA, LB, LC, LD, DROW>::value))) void
jk_group_kernel(const JKGroup *__restrict__ groups, int ngroup, const JKTask *__restrict__ tasks,
                int *__restrict__ counter, const unsigned char *__restrict__ flags,
                const unsigned char *__restrict__ gmask, int accept, DevBasis B, const double *__restrict__ D,
                double *__restrict__ M, int nao, double *__restrict__ jslab, int tfirst,
                const PrimPair *__restrict__ pp, MixArgs mx)
{
    using C = Class<LA, LB, LC, LD>;
    using R64 = RysLds<LA, LB, LC, LD, double>;
    using R32 = RysLds<LA, LB, LC, LD, float>;
    // the primitive-pair records through a restrict-qualified argument:
    // read-only and unaliased, so wavefront-uniform (bra) records can be
    // fetched with scalar loads
    B.pp = pp;
    extern __shared__ __align__(16) double lds[];
    double *rt = lds;                                                        // FP64 Rys table block
    float *rt32 = (float *)(lds + RysLdsDoubles<LA, LB, LC, LD, double>()); // FP32 block (MIXED)
    double *mab_l = lds + GroupTabDoubles<LA, LB, LC, LD, MIXED>();         // bra-bra block
    double *Krow = mab_l + C::NFA * C::NFB;                                 // (NFA + NFB) rows of nao
    double *Drow = Krow + (C::NFA + C::NFB) * nao;                          // density rows (DROW)
    __shared__ int g_s;
    rys_stage<double, R64::N, R64::WONLY>(rt);
    if constexpr (MIXED) rys_stage<float, R32::N, R32::WONLY>(rt32);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nrow = (C::NFA + C::NFB) * nao;
    const size_t n = (size_t)nao;
    for (;;) {
        if (threadIdx.x == 0) {
            int g;
            do {
                g = atomicAdd(counter, 1);
            } while (g >= 0 && g < ngroup && gmask && !(gmask[g] & accept));
            g_s = g;
        }
        for (int i = threadIdx.x; i < nrow; i += blockDim.x) Krow[i] = 0.0;
        if (threadIdx.x < C::NFA * C::NFB) mab_l[threadIdx.x] = 0.0;
        __syncthreads();
        // the value is the same in every lane; readfirstlane tells the compiler
        // so, which turns the bra loads below into scalar loads
        const int g = __builtin_amdgcn_readfirstlane(g_s);
        if (g >= ngroup || g < 0) break;
        const JKGroup G = groups[g];
        const PairView bra = load_pair(B, G.bra);
        if constexpr (DROW) {
            for (int i = threadIdx.x; i < nrow; i += blockDim.x) {
                const int r = i / nao, col = i - r * nao;
                const size_t row = (r < C::NFA) ? (size_t)(bra.aoa + r) : (size_t)(bra.aob + r - C::NFA);
                Drow[i] = D[row * n + col];
            }
            __syncthreads();
        }
        double mab[C::NFA * C::NFB];
#pragma unroll
        for (int e = 0; e < C::NFA * C::NFB; e++) mab[e] = 0.0;
        // this wavefront's tasks of the group are G.first + wave + NW j (j <
        // 64: groups hold at most 256 tasks); lane j holds the flag of the j-th
        int myflag = 0;
        constexpr int NW = GroupNW<LA, LB, LC, LD, DROW>::value;
        if (flags && G.first + wave + NW * lane < G.first + G.count) myflag = flags[G.first + wave + NW * lane];
        double qdb = 0.0, qb = 0.0;
        if constexpr (MIXED) {
            qb = B.pair_rb[G.bra].w;
            qdb = mx.Dpair[G.bra];
        }
        for (int t = G.first + wave, j = 0; t < G.first + G.count; t += NW, j++) {
            const int fl = flags ? __builtin_amdgcn_readlane(myflag, j) : 0;
            if (!((accept >> fl) & 1)) continue;
            const JKTask tk = tasks[t];
            double *js = jslab ? jslab + (size_t)(t - tfirst) * (64 * C::NFC * C::NFD) : nullptr;
            bool f32 = false;
            if constexpr (MIXED) {
                const int k0 = __builtin_amdgcn_readfirstlane(tk.ket0);
                f32 = qb * fmax(qdb * B.pair_rb[k0].w, mx.Wqd[k0]) < mx.tau_mix;
                if (lane == 0) mx.prec[t] = f32 ? 1 : 0;
            }
            if (f32)
                jk_task_rows<LA, LB, LC, LD, float, DROW>(tk, bra, B, D, M, nao, lane, rt32, Krow, mab, js, Drow,
                                                          mab_l);
            else
                jk_task_rows<LA, LB, LC, LD, double, DROW>(tk, bra, B, D, M, nao, lane, rt, Krow, mab, js, Drow,
                                                           mab_l);
        }
        if constexpr (Split<LA, LB, LC, LD>::N == 1) {
#pragma unroll
            for (int e = 0; e < C::NFA * C::NFB; e++) {
                const double v = wave_sum(mab[e]);
                if (lane == 0) atomicAdd(&mab_l[e], v);
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nrow; i += blockDim.x) {
            const double v = Krow[i];
            if (v != 0.0) {
                const int r = i / nao, col = i - r * nao;
                const size_t row = (r < C::NFA) ? (size_t)(bra.aoa + r) : (size_t)(bra.aob + r - C::NFA);
                atomicAdd(&M[row * n + col], v);
            }
        }
        if (threadIdx.x < C::NFA * C::NFB)
            atomicAdd(&M[(bra.aoa + threadIdx.x / C::NFB) * n + bra.aob + threadIdx.x % C::NFB], mab_l[threadIdx.x]);
        __syncthreads();
    }
}

template <int LA, int LB, int LC, int LD, bool MIXED>
static size_t group_lds_bytes(int nao, bool drow)
{
    using C = Class<LA, LB, LC, LD>;
    return sizeof(double) * ((size_t)GroupTabDoubles<LA, LB, LC, LD, MIXED>() + C::NFA * C::NFB +
                             (size_t)(C::NFA + C::NFB) * nao * (drow ? 2 : 1));
}

// Persistent-grid size (resident workgroups per CU x CUs) of a kernel at a
// given dynamic LDS size on the current device, computed once per (kernel,
// device, LDS size) under a lock, so concurrent callers and several devices
// each get their own.
static int persistent_grid(const void *kernel, int threads, size_t lds, const char *what)
{
    static std::mutex mu;
    static std::map<std::tuple<const void *, int, size_t>, int> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const auto key = std::make_tuple(kernel, dev, lds);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int ncu = 0, per = 0;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, lds);
    const int grid = std::max(1, ncu) * std::max(1, per);
    cache[key] = grid;
    if (getenv("LIBRINT_DEBUG"))
        fprintf(stderr, "librint: %s lds %zu blocks/cu %d grid %d (device %d)\n", what, lds, per, grid, dev);
    return grid;
}

template <int LA, int LB, int LC, int LD, bool MIXED, bool DROW>
static hipError_t launch_jk_group_d(const JKGroup *groups, int ngroup, const JKTask *tasks, int *counter,
                                    const unsigned char *flags, const unsigned char *gmask, int accept,
                                    const DevBasis &B, const double *D, double *M, int nao, double *jslab,
                                    int tfirst, const MixArgs &mx, hipStream_t st)
{
    const size_t lds = group_lds_bytes<LA, LB, LC, LD, MIXED>(nao, DROW);
    char what[64];
    snprintf(what, sizeof what, "jk_group<%d%d|%d%d>%s%s", LA, LB, LC, LD, MIXED ? " mixed" : "", DROW ? " drow" : "");
    constexpr int threads = 64 * GroupNW<LA, LB, LC, LD, DROW>::value;
    const int grid = persistent_grid((const void *)jk_group_kernel<LA, LB, LC, LD, MIXED, DROW>, threads, lds, what);
    const int g = std::min(grid, ngroup);
    hipLaunchKernelGGL((jk_group_kernel<LA, LB, LC, LD, MIXED, DROW>), dim3(g), dim3(threads), lds, st, groups, ngroup,
                       tasks, counter, flags, gmask, accept, B, D, M, nao, jslab, tfirst, B.pp, mx);
    return hipGetLastError();
}

template <int LA, int LB, int LC, int LD, bool MIXED>
static hipError_t launch_jk_group_t(const JKGroup *groups, int ngroup, const JKTask *tasks, int *counter,
                                    const unsigned char *flags, const unsigned char *gmask, int accept,
                                    const DevBasis &B, const double *D, double *M, int nao, double *jslab,
                                    int tfirst, bool drow, const MixArgs &mx, hipStream_t st)
{
    return drow ? launch_jk_group_d<LA, LB, LC, LD, MIXED, true>(groups, ngroup, tasks, counter, flags, gmask, accept,
                                                                  B, D, M, nao, jslab, tfirst, mx, st)
                : launch_jk_group_d<LA, LB, LC, LD, MIXED, false>(groups, ngroup, tasks, counter, flags, gmask,
                                                                   accept, B, D, M, nao, jslab, tfirst, mx, st);
}

// ------------------------------------------------------ team J/K kernel --
// Classes whose contracted block does not fit in one lane's registers
// (NF > 108, e.g. (dd|dd) with 1296 integrals and 5 roots): one wavefront
// works on one quartet at a time, looping over the kets of its task.
// Per primitive quartet, lanes c < 3*NROOTS build one (root, axis) chain of
// the VRR/HRR in LDS; then every lane accumulates its share of the NF
// outputs (f = lane + 64k) from the LDS chains.  The J/K blocks of a quartet
// are reduced in LDS (ds_add_f64) and flushed with one global atomic per
// element; the bra-bra block stays in LDS for the whole task.
__device__ __constant__ signed char CARTP[3][6][3] = {
    {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}}};


template <int LA, int LB, int LC, int LD>
struct TeamShape {
    using C = Class<LA, LB, LC, LD>;
    static constexpr int NCH = 3 * C::NROOTS;
    static constexpr int NPL = (C::NF + 63) / 64;
    static constexpr int NAB = C::NFA * C::NFB, NCD = C::NFC * C::NFD;
    static constexpr int NAC = C::NFA * C::NFC, NBC = C::NFB * C::NFC;
    static constexpr int NAD = C::NFA * C::NFD, NBD = C::NFB * C::NFD;
    static constexpr int NKET = NCD + NAC + NBC + NAD + NBD;
};

// one primitive quartet's chains into LDS (lanes c < NCH)
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void team_chain(double *g, int c, const PrimPair &p, const PrimPair &q, const double rbd[3],
                                           const double rab[3], const double rcd[3])
{
    using C = Class<LA, LB, LC, LD>;
    const double aij = p.e, akl = q.e;
    const double rsq = rsqrt_pos(aij + akl);
    const double rs = rsq * rsq;
    const double a0 = aij * akl * rs;
    const double pqx = p.pbx - q.pbx + rbd[0], pqy = p.pby - q.pby + rbd[1], pqz = p.pbz - q.pbz + rbd[2];
    const double rr = pqx * pqx + pqy * pqy + pqz * pqz;
    const double fac1 = p.K * q.K * rsq;
    const int r = c / 3, ax = c - 3 * (c / 3);
    double t2, wr;
    rys_one<C::NROOTS>(a0 * rr, r, t2, wr);  // only this lane's root
    const double tr = t2 * rs;
    const double b00 = 0.5 * tr;
    const double b10 = p.h * (1.0 - akl * tr);
    const double b01 = q.h * (1.0 - aij * tr);
    const double pq = sel3(ax, pqx, pqy, pqz);
    const double cn = sel3(ax, p.pbx, p.pby, p.pbz) - akl * tr * pq;
    const double cm = sel3(ax, q.pbx, q.pby, q.pbz) + aij * tr * pq;
    double *gc = g + c * C::GS;
    gc[0] = (ax == 2) ? wr * fac1 : 1.0;
    vrr_axis<C::NMAX, C::MMAX, C::DL, C::DJ>(gc, cn, cm, b10, b01, b00);
    if constexpr (LA > 0 || LC > 0)
        hrr_axis<LA, LB, LC, C::NMAX, C::MMAX, C::DK, C::DL, C::DJ>(gc, sel3(ax, rab[0], rab[1], rab[2]),
                                                                   sel3(ax, rcd[0], rcd[1], rcd[2]));
}

// contracted block of one quartet, distributed: acc[k] holds f = lane + 64k
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void team_contracted(double *acc, double *g, const int *pk, const DevBasis &B,
                                                const PairView &bra, const PairView &ket, int lane)
{
    using C = Class<LA, LB, LC, LD>;
    using S = TeamShape<LA, LB, LC, LD>;
#pragma unroll
    for (int k = 0; k < S::NPL; k++) acc[k] = 0.0;
    const double rbd[3] = {bra.rb[0] - ket.rb[0], bra.rb[1] - ket.rb[1], bra.rb[2] - ket.rb[2]};
    for (int ib = 0; ib < bra.npp; ib++) {
        const PrimPair p = B.pp[bra.off + ib];
        for (int ik = 0; ik < ket.npp; ik++) {
            const PrimPair q = B.pp[ket.off + ik];
            if (lane < S::NCH) team_chain<LA, LB, LC, LD>(g, lane, p, q, rbd, bra.rab, ket.rab);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < S::NPL; k++) {
                if (lane + 64 * k >= C::NF) continue;
                const int ix = pk[k] & 255, iy = (pk[k] >> 8) & 255, iz = pk[k] >> 16;
                double a = acc[k];
#pragma unroll
                for (int r = 0; r < C::NROOTS; r++)
                    a += g[(3 * r) * C::GS + ix] * g[(3 * r + 1) * C::GS + iy] * g[(3 * r + 2) * C::GS + iz];
                acc[k] = a;
            }
            __syncthreads();
        }
    }
}

// packed g offsets (ix | iy << 8 | iz << 16, all < GS <= 225) of the outputs
// f = lane + 64k a lane owns
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void team_indices(int lane, int *pk)
{
    using C = Class<LA, LB, LC, LD>;
    using S = TeamShape<LA, LB, LC, LD>;
    static_assert(C::GS <= 256, "packed offsets");
#pragma unroll
    for (int k = 0; k < S::NPL; k++) {
        const int f = lane + 64 * k;
        const int ff = f < C::NF ? f : 0;
        const int fa = ff % C::NFA, t1 = ff / C::NFA;
        const int fb = t1 % C::NFB, t2 = t1 / C::NFB;
        const int fc = t2 % C::NFC, fd = t2 / C::NFC;
        const int ix = CARTP[LB][fb][0] * C::DJ + CARTP[LD][fd][0] * C::DL + CARTP[LC][fc][0] * C::DK + CARTP[LA][fa][0];
        const int iy = CARTP[LB][fb][1] * C::DJ + CARTP[LD][fd][1] * C::DL + CARTP[LC][fc][1] * C::DK + CARTP[LA][fa][1];
        const int iz = CARTP[LB][fb][2] * C::DJ + CARTP[LD][fd][2] * C::DL + CARTP[LC][fc][2] * C::DK + CARTP[LA][fa][2];
        pk[k] = ix | (iy << 8) | (iz << 16);
    }
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void jk_team_kernel(const JKTask *__restrict__ tasks,
                                                     const unsigned char *__restrict__ flags, int accept, DevBasis B,
                                                     const double *__restrict__ D, double *__restrict__ M, int nao)
{
    using C = Class<LA, LB, LC, LD>;
    using S = TeamShape<LA, LB, LC, LD>;
    __shared__ double g[S::NCH * C::GS];
    __shared__ double accs[S::NPL * 64];  // contracted block V[f], f = fa + NFA (fb + NFB (fc + NFC fd))
    // density blocks: D_ab of the bra (per task), then D_cd, D_ac, D_bc,
    // D_ad, D_bd of the ket (the order of the ket-side outputs below)
    __shared__ double dl[S::NAB + S::NKET];
    if (flags && !((accept >> flags[blockIdx.x]) & 1)) return;
    const JKTask t = tasks[blockIdx.x];
    const int lane = threadIdx.x;
    const PairView bra = load_pair(B, t.bra);
    int pk[S::NPL];
    team_indices<LA, LB, LC, LD>(lane, pk);
    const size_t n = (size_t)nao;
    constexpr int SC = C::NFA * C::NFB, SD = SC * C::NFC;  // strides of fc, fd in V
    constexpr int OCD = S::NAB, OAC = OCD + S::NCD, OBC = OAC + S::NAC, OAD = OBC + S::NBC, OBD = OAD + S::NAD;
    constexpr int NKPL = (S::NKET + 63) / 64;
    for (int e = lane; e < S::NAB; e += 64) dl[e] = D[(bra.aoa + e / C::NFB) * n + bra.aob + e % C::NFB];
    double jab = 0.0;  // lane < NAB: Coulomb element (fa fb) of the bra, summed over the kets
    for (int kk = 0; kk < t.nket; kk++) {
        const int ketp = t.ket0 + kk;
        const PairView ket = load_pair(B, ketp);
        double acc[S::NPL];
        team_contracted<LA, LB, LC, LD>(acc, g, pk, B, bra, ket, lane);
        double s = 1.0;
        if (bra.sa == bra.sb) s *= 0.5;
        if (ket.sa == ket.sb) s *= 0.5;
        if (t.bra == ketp) s *= 0.5;
#pragma unroll
        for (int k = 0; k < S::NPL; k++) accs[k * 64 + lane] = acc[k];
        // ket-side element e: its M target (row, col) is also where its
        // density block entry comes from (M[c][d] <- D[c][d], M[a][c] <- D[a][c], ...)
        size_t rc[NKPL];
#pragma unroll
        for (int j = 0; j < NKPL; j++) {
            const int e = lane + 64 * j;
            rc[j] = 0;
            if (e >= S::NKET) continue;
            size_t row, col;
            if (e < S::NCD) {
                row = ket.aoa + e / C::NFD;
                col = ket.aob + e % C::NFD;
            } else if (e < S::NCD + S::NAC) {
                const int x = e - S::NCD;
                row = bra.aoa + x / C::NFC;
                col = ket.aoa + x % C::NFC;
            } else if (e < S::NCD + S::NAC + S::NBC) {
                const int x = e - S::NCD - S::NAC;
                row = bra.aob + x / C::NFC;
                col = ket.aoa + x % C::NFC;
            } else if (e < S::NCD + S::NAC + S::NBC + S::NAD) {
                const int x = e - S::NCD - S::NAC - S::NBC;
                row = bra.aoa + x / C::NFD;
                col = ket.aob + x % C::NFD;
            } else {
                const int x = e - S::NCD - S::NAC - S::NBC - S::NAD;
                row = bra.aob + x / C::NFD;
                col = ket.aob + x % C::NFD;
            }
            rc[j] = row * n + col;
            dl[S::NAB + e] = D[rc[j]];
        }
        __syncthreads();
        // every J/K element is a small contraction of V with one density block
        if (lane < S::NAB) {
            const int fa = lane / C::NFB, fb = lane - fa * C::NFB;
            double v = 0.0;
#pragma unroll
            for (int fd = 0; fd < C::NFD; fd++)
#pragma unroll
                for (int fc = 0; fc < C::NFC; fc++)
                    v += accs[fa + C::NFA * fb + SC * fc + SD * fd] * dl[OCD + fc * C::NFD + fd];
            jab += 2.0 * s * v;
        }
#pragma unroll
        for (int j = 0; j < NKPL; j++) {
            const int e = lane + 64 * j;
            if (e >= S::NKET) continue;
            // summed V indices i1 < n1 (stride vs1), i2 < n2 (stride vs2)
            // against dl[db + i1*ds1 + i2]
            int vb, vs1, vs2, n1, n2, db, ds1;
            double coef = -0.5;
            if (e < S::NCD) {  // M[c][d] += 2 s sum_ab V D_ab
                const int fc = e / C::NFD, fd = e - fc * C::NFD;
                vb = SC * fc + SD * fd, vs1 = 1, vs2 = C::NFA, n1 = C::NFA, n2 = C::NFB, db = 0, ds1 = C::NFB;
                coef = 2.0;
            } else if (e < S::NCD + S::NAC) {  // M[a][c] -= s/2 sum_bd V D_bd
                const int x = e - S::NCD, fa = x / C::NFC, fc = x - fa * C::NFC;
                vb = fa + SC * fc, vs1 = C::NFA, vs2 = SD, n1 = C::NFB, n2 = C::NFD, db = OBD, ds1 = C::NFD;
            } else if (e < S::NCD + S::NAC + S::NBC) {  // M[b][c] -= s/2 sum_ad V D_ad
                const int x = e - S::NCD - S::NAC, fb = x / C::NFC, fc = x - fb * C::NFC;
                vb = C::NFA * fb + SC * fc, vs1 = 1, vs2 = SD, n1 = C::NFA, n2 = C::NFD, db = OAD, ds1 = C::NFD;
            } else if (e < S::NCD + S::NAC + S::NBC + S::NAD) {  // M[a][d] -= s/2 sum_bc V D_bc
                const int x = e - S::NCD - S::NAC - S::NBC, fa = x / C::NFD, fd = x - fa * C::NFD;
                vb = fa + SD * fd, vs1 = C::NFA, vs2 = SC, n1 = C::NFB, n2 = C::NFC, db = OBC, ds1 = C::NFC;
            } else {  // M[b][d] -= s/2 sum_ac V D_ac
                const int x = e - S::NCD - S::NAC - S::NBC - S::NAD, fb = x / C::NFD, fd = x - fb * C::NFD;
                vb = C::NFA * fb + SD * fd, vs1 = 1, vs2 = SC, n1 = C::NFA, n2 = C::NFC, db = OAC, ds1 = C::NFC;
            }
            double v = 0.0;
            for (int i1 = 0; i1 < n1; i1++)
                for (int i2 = 0; i2 < n2; i2++) v += accs[vb + vs1 * i1 + vs2 * i2] * dl[db + ds1 * i1 + i2];
            atomicAdd(&M[rc[j]], coef * s * v);
        }
        __syncthreads();
    }
    if (lane < S::NAB) atomicAdd(&M[(bra.aoa + lane / C::NFB) * n + bra.aob + lane % C::NFB], jab);
}

// list variant for the team classes: one wavefront per list task
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void eri_list_team_kernel(const ListTask *__restrict__ tasks, int ntask, DevBasis B,
                                                           double *__restrict__ out, int mode, int nao)
{
    using C = Class<LA, LB, LC, LD>;
    using S = TeamShape<LA, LB, LC, LD>;
    __shared__ double g[S::NCH * C::GS];
    __shared__ double red[64];
    if ((int)blockIdx.x >= ntask) return;
    const ListTask t = tasks[blockIdx.x];
    const int lane = threadIdx.x;
    int pk[S::NPL];
    team_indices<LA, LB, LC, LD>(lane, pk);
    const PairView bra = load_pair(B, t.bra);
    const PairView ket = load_pair(B, t.ket);
    double acc[S::NPL];
    team_contracted<LA, LB, LC, LD>(acc, g, pk, B, bra, ket, lane);
    const size_t n = (size_t)nao;
    if (mode == 0) {
#pragma unroll
        for (int k = 0; k < S::NPL; k++)
            if (lane + 64 * k < C::NF) out[t.out + lane + 64 * k] = acc[k];
    } else if (mode == 1) {
        double m = 0.0;
        if constexpr (LA == LC && LB == LD) {
#pragma unroll
            for (int k = 0; k < S::NPL; k++) {
                const int f = lane + 64 * k;
                if (f >= C::NF) continue;
                const int fa = f % C::NFA, t1 = f / C::NFA;
                const int fb = t1 % C::NFB, t2 = t1 / C::NFB;
                const int fc = t2 % C::NFC, fd = t2 / C::NFC;
                if (fa == fc && fb == fd) m = fmax(m, fabs(acc[k]));
            }
        }
        red[lane] = m;
        __syncthreads();
        if (lane == 0) {
            double mm = 0.0;
            for (int e = 0; e < 64; e++) mm = fmax(mm, red[e]);
            out[t.out] = sqrt(mm);
        }
    } else {
        // entries related by a same-shell pair or by bra = ket map to the
        // same 8 addresses and are held by different lanes: only the
        // canonical one writes, so every image carries one value
        const bool sab = bra.sa == bra.sb, scd = ket.sa == ket.sb, sbk = t.bra == t.ket;
#pragma unroll
        for (int k = 0; k < S::NPL; k++) {
            const int f = lane + 64 * k;
            if (f >= C::NF) continue;
            const int fa = f % C::NFA, t1 = f / C::NFA;
            const int fb = t1 % C::NFB, t2 = t1 / C::NFB;
            const int fc = t2 % C::NFC, fd = t2 / C::NFC;
            if ((sab && fa < fb) || (scd && fc < fd) || (sbk && fa * C::NFB + fb < fc * C::NFD + fd)) continue;
            const double v = acc[k];
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

// Shell-pair block maxima of |D|: Dsh[i*nsh + j] = max |D[mu][nu]|, mu in
// shell i, nu in shell j (virtual segmented shells).
__global__ void dsh_kernel(const double *__restrict__ D, const int *__restrict__ ao_off, int nsh, int nao,
                           double *__restrict__ Dsh)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nsh * nsh) return;
    const int i = (int)(idx / nsh), j = (int)(idx % nsh);
    const int i0 = ao_off[i], i1 = ao_off[i + 1], j0 = ao_off[j], j1 = ao_off[j + 1];
    double m = 0.0;
    for (int a = i0; a < i1; a++)
        for (int b = j0; b < j1; b++) m = fmax(m, fabs(D[(size_t)a * nao + b]));
    Dsh[idx] = m;
}

// One wavefront per task: the density-weighted Schwarz bound of each quartet
// Q_ab Q_cd max(|D| over the six blocks the quartet reads) decides the task
// flag from the largest bound of its kets: < tau_skip -> 2 (skip),
// < tau_mix -> 1 (FP32 recurrences), else 0 (FP64).  gmask[group] collects
// the flags present in each group.
__global__ __launch_bounds__(256) void classify_kernel(const JKTask *__restrict__ tasks, int ntask,
                                                       const int *__restrict__ task_group, DevBasis B,
                                                       const double *__restrict__ Dsh, int nsh, double tau_mix,
                                                       double tau_skip, unsigned char *__restrict__ flags,
                                                       unsigned int *__restrict__ gmask)
{
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= ntask) return;
    const JKTask T = tasks[t];
    const int4 bi = B.pair_info[T.bra];
    const double qb = B.pair_rb[T.bra].w;
    double bound = 0.0;
    if (lane < T.nket) {
        const int4 ki = B.pair_info[T.ket0 + lane];
        const double qk = B.pair_rb[T.ket0 + lane].w;
        const size_t a = bi.x, b = bi.y, c = ki.x, d = ki.y, n = nsh;
        double dm = fmax(fmax(Dsh[a * n + b], Dsh[c * n + d]), fmax(Dsh[a * n + c], Dsh[a * n + d]));
        dm = fmax(dm, fmax(Dsh[b * n + c], Dsh[b * n + d]));
        bound = qb * qk * dm;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) bound = fmax(bound, __shfl_xor(bound, off, 64));
    if (lane == 0) {
        const unsigned char f = bound < tau_skip ? 2 : (bound < tau_mix ? 1 : 0);
        flags[t] = f;
        atomicOr(&gmask[task_group[t]], 1u << f);
    }
}

__global__ void zero_u32_kernel(unsigned int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// gmask words -> bytes for the group kernels
__global__ void pack_gmask_kernel(const unsigned int *__restrict__ w, unsigned char *__restrict__ b, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (unsigned char)w[i];
}

// Mixed-precision bounds per build (from Dsh):
//   Drow[s]  = max_j Dsh[s][j]
//   Dpair[p] = max(Drow[sa], Drow[sb])   >= |D| over every block a quartet
//              with pair p reads on p's side (ab, ac, ad, bc, bd)
//   Wqd[k]   = max over k' in [k, k + 64) of Q[k'] Dpair[k']
// so that a task (bra b, kets [ket0, ket0 + nket), nket <= 64, Q descending)
// has every quartet bound Q_b Q_k max(Dsh of its six blocks) below
//   Q_b max(Dpair[b] Q[ket0], Wqd[ket0]).
__global__ void drow_kernel(const double *__restrict__ Dsh, int nsh, double *__restrict__ Drow)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsh) return;
    double m = 0.0;
    for (int j = 0; j < nsh; j++) m = fmax(m, Dsh[(size_t)s * nsh + j]);
    Drow[s] = m;
}

__global__ void dpair_kernel(const int4 *__restrict__ pair_info, const double4 *__restrict__ pair_rb, int npair,
                             const double *__restrict__ Drow, double *__restrict__ Dpair, double *__restrict__ qd)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npair) return;
    const int4 pi = pair_info[p];
    const double d = fmax(Drow[pi.x], Drow[pi.y]);
    Dpair[p] = d;
    qd[p] = d * pair_rb[p].w;
}

__global__ void wqd_kernel(const double *__restrict__ qd, int npair, double *__restrict__ Wqd)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npair) return;
    double m = 0.0;
    const int e = min(npair, k + 64);
    for (int j = k; j < e; j++) m = fmax(m, qd[j]);
    Wqd[k] = m;
}

hipError_t launch_mix_bounds(const double *Dsh, int nsh, const DevBasis &B, int npair, double *Drow, double *Dpair,
                             double *qd, double *Wqd, hipStream_t st)
{
    if (nsh <= 0 || npair <= 0) return hipSuccess;
    hipLaunchKernelGGL(drow_kernel, dim3((nsh + 255) / 256), dim3(256), 0, st, Dsh, nsh, Drow);
    hipLaunchKernelGGL(dpair_kernel, dim3((npair + 255) / 256), dim3(256), 0, st, B.pair_info, B.pair_rb, npair, Drow,
                       Dpair, qd);
    hipLaunchKernelGGL(wqd_kernel, dim3((npair + 255) / 256), dim3(256), 0, st, qd, npair, Wqd);
    return hipGetLastError();
}

hipError_t launch_dsh(const double *D, const int *ao_off, int nsh, int nao, double *Dsh, hipStream_t st)
{
    const long long nn = (long long)nsh * nsh;
    hipLaunchKernelGGL(dsh_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, D, ao_off, nsh, nao, Dsh);
    return hipGetLastError();
}

hipError_t launch_classify(const JKTask *tasks, int ntask, const int *task_group, int ngroup, const DevBasis &B,
                           const double *D, const int *ao_off, int nsh, int nao, double *Dsh, double tau_mix,
                           double tau_skip, unsigned char *flags, unsigned int *gmask_w, unsigned char *gmask,
                           hipStream_t st)
{
    (void)D;
    (void)ao_off;
    (void)nao;
    if (ngroup > 0)
        hipLaunchKernelGGL(zero_u32_kernel, dim3((ngroup + 255) / 256), dim3(256), 0, st, gmask_w, ngroup);
    if (ntask > 0)
        hipLaunchKernelGGL(classify_kernel, dim3((ntask + 3) / 4), dim3(256), 0, st, tasks, ntask, task_group, B, Dsh,
                           nsh, tau_mix, tau_skip, flags, gmask_w);
    if (ngroup > 0)
        hipLaunchKernelGGL(pack_gmask_kernel, dim3((ngroup + 255) / 256), dim3(256), 0, st, gmask_w, gmask, ngroup);
    return hipGetLastError();
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
// Pair classes (la <= lb): 0:(0,0) 1:(0,1) 2:(1,1) 3:(0,2) 4:(1,2) 5:(2,2);
// bra class >= ket class.  REG: contracted block in registers (NF <= 108);
// TEAM: one wavefront per quartet.
#define LR_REG_CLASSES(X)                                                                                              \
    X(0, 0, 0, 0) X(0, 1, 0, 0) X(0, 1, 0, 1) X(1, 1, 0, 0) X(1, 1, 0, 1) X(1, 1, 1, 1) X(0, 2, 0, 0) X(0, 2, 0, 1)   \
        X(0, 2, 1, 1) X(0, 2, 0, 2) X(1, 2, 0, 0) X(1, 2, 0, 1) X(1, 2, 0, 2) X(2, 2, 0, 0) X(2, 2, 0, 1)
// SPLIT: contracted block in registers chunk by chunk (Split<>), team
// kernel when the group kernel's LDS rows do not fit
#define LR_SPLIT_CLASSES(X) X(1, 2, 1, 1) X(1, 2, 1, 2) X(2, 2, 1, 1) X(2, 2, 0, 2)
#define LR_TEAM_CLASSES(X) X(2, 2, 1, 2) X(2, 2, 2, 2)
// the list kernel keeps fewer live values than the J/K kernel but no fewer
// outputs; these two also go to the team form there
#define LR_LIST_REG_CLASSES(X)                                                                                         \
    X(0, 0, 0, 0) X(0, 1, 0, 0) X(0, 1, 0, 1) X(1, 1, 0, 0) X(1, 1, 0, 1) X(1, 1, 1, 1) X(0, 2, 0, 0) X(0, 2, 0, 1)   \
        X(0, 2, 1, 1) X(0, 2, 0, 2) X(1, 2, 0, 0) X(1, 2, 0, 1) X(2, 2, 0, 0)
#define LR_LIST_TEAM_CLASSES(X) LR_SPLIT_CLASSES(X) LR_TEAM_CLASSES(X) X(1, 2, 0, 2) X(2, 2, 0, 1)

bool class_supported(int la, int lb, int lc, int ld)
{
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d) return true;
    LR_REG_CLASSES(X)
    LR_SPLIT_CLASSES(X)
    LR_TEAM_CLASSES(X)
#undef X
    return false;
}

bool class_split(int la, int lb, int lc, int ld)
{
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d) return true;
    LR_SPLIT_CLASSES(X)
#undef X
    return false;
}

template <int LA, int LB, int LC, int LD, typename T>
static hipError_t launch_jk_persistent(const JKTask *tasks, int ntask, int *counter, const unsigned char *flags,
                                       int accept, const DevBasis &B, const double *D, double *M, int nao,
                                       double *jslab, int tfirst, hipStream_t st)
{
    char what[48];
    snprintf(what, sizeof what, "jk<%d%d|%d%d>", LA, LB, LC, LD);
    const int grid = persistent_grid((const void *)jk_kernel<LA, LB, LC, LD, T>, 256, 0, what);
    const int g = std::min(grid, (ntask + 3) / 4);
    hipLaunchKernelGGL((jk_kernel<LA, LB, LC, LD, T>), dim3(g), dim3(256), 0, st, tasks, ntask, counter, flags, accept, B,
                       D, M, nao, jslab, tfirst);
    return hipGetLastError();
}

__global__ void zero_ints_kernel(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void zero_doubles_kernel(double *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

hipError_t launch_zero_doubles(double *p, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_doubles_kernel, dim3(blocks), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_zero_counters(int *p, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

bool class_has_fp32(int la, int lb, int lc, int ld)
{
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d) return true;
    LR_REG_CLASSES(X)
    LR_SPLIT_CLASSES(X)
#undef X
    return false;
}

hipError_t launch_jk(int la, int lb, int lc, int ld, int fp32, const JKTask *tasks, int ntask, int *counter,
                     const unsigned char *flags, int accept, const DevBasis &B, const double *D, double *M, int nao,
                     double *jslab, int tfirst, hipStream_t st)
{
    if (ntask == 0) return hipSuccess;
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d)                                                                      \
        return launch_jk_persistent<a, b, c, d, double>(tasks, ntask, counter, flags, accept, B, D, M, nao, jslab,     \
                                                        tfirst, st);
    LR_REG_CLASSES(X)
#undef X
    if (fp32 || jslab) return hipErrorInvalidValue;
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d) {                                                                    \
        hipLaunchKernelGGL((jk_team_kernel<a, b, c, d>), dim3(ntask), dim3(64), 0, st, tasks, flags, accept, B, D, M,   \
                           nao);                                                                                       \
        return hipGetLastError();                                                                                      \
    }
    LR_SPLIT_CLASSES(X)
    LR_TEAM_CLASSES(X)
#undef X
    return hipErrorInvalidValue;
}

int jk_group_waves(int la, int lb, int lc, int ld, bool drow)
{
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d)                                                                      \
        return drow ? GroupNW<a, b, c, d, true>::value : GroupNW<a, b, c, d, false>::value;
    LR_REG_CLASSES(X)
    LR_SPLIT_CLASSES(X)
#undef X
    return 4;
}

size_t jk_group_lds(int la, int lb, int lc, int ld, int mixed, int nao, bool drow)
{
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d)                                                                      \
        return mixed ? group_lds_bytes<a, b, c, d, true>(nao, drow) : group_lds_bytes<a, b, c, d, false>(nao, drow);
    LR_REG_CLASSES(X)
    LR_SPLIT_CLASSES(X)
#undef X
    return (size_t)-1;
}

hipError_t launch_jk_group(int la, int lb, int lc, int ld, int mixed, const JKGroup *groups, int ngroup,
                           const JKTask *tasks, int *counter, const unsigned char *flags, const unsigned char *gmask,
                           int accept, const DevBasis &B, const double *D, double *M, int nao, double *jslab,
                           int tfirst, bool drow, const MixArgs &mx, hipStream_t st)
{
    if (ngroup == 0) return hipSuccess;
    if (mixed && (!mx.Dpair || !mx.Wqd || !mx.prec)) return hipErrorInvalidValue;
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d)                                                                      \
        return mixed ? launch_jk_group_t<a, b, c, d, true>(groups, ngroup, tasks, counter, flags, gmask, accept, B, D,  \
                                                           M, nao, jslab, tfirst, drow, mx, st)                        \
                     : launch_jk_group_t<a, b, c, d, false>(groups, ngroup, tasks, counter, flags, gmask, accept, B,   \
                                                            D, M, nao, jslab, tfirst, drow, mx, st);
    LR_REG_CLASSES(X)
    LR_SPLIT_CLASSES(X)
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
    LR_LIST_REG_CLASSES(X)
#undef X
#define X(a, b, c, d)                                                                                                  \
    if (la == a && lb == b && lc == c && ld == d) {                                                                    \
        hipLaunchKernelGGL((eri_list_team_kernel<a, b, c, d>), dim3(ntask), dim3(64), 0, st, tasks, ntask, B, out,    \
                           mode, nao);                                                                                 \
        return hipGetLastError();                                                                                      \
    }
    LR_LIST_TEAM_CLASSES(X)
#undef X
    return hipErrorInvalidValue;
}

// One workgroup of 4 wavefronts per (item, ket component e): lane = ket offset
// in the window; the wavefronts split the item's tasks, each summing the
// slab entries of the executed tasks whose ket range covers the lane; the
// partial sums are combined in LDS and added to M[c][d] with one atomic.
// Each wavefront issues the slab loads of 8 tasks before adding them, so the
// reduction runs on memory bandwidth, not on one load latency per task.
__global__ __launch_bounds__(256) void jreduce_kernel(const JRedItem *__restrict__ items, int nitem,
                                                      const int *__restrict__ item_tasks,
                                                      const JKTask *__restrict__ tasks,
                                                      const unsigned char *__restrict__ flags, int accept,
                                                      const double *__restrict__ J, DevBasis B,
                                                      double *__restrict__ M, int nao)
{
    __shared__ double part[4][64];
    __shared__ int seen[4][64];
    const int ii = blockIdx.x;
    if (ii >= nitem) return;
    const JRedItem it = items[ii];
    const int nfcd = it.nfc * it.nfd;
    const int e = blockIdx.y;
    if (e >= nfcd) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc = 0.0;
    int any = 0;
    const double *Je = J + it.jbase + e * 64 + lane;
    const long long stride = 64LL * nfcd;
    // wave w takes the item's tasks k = w, w + 4, ...: their indices, flags
    // and ket counts are loaded at once (lane j holds task w + 4 j of the
    // current block of 256), so the slab loads below do not wait on them
    for (int base = wave; base < it.lcount; base += 256) {
        const int kk = base + 4 * lane;
        int tl = 0, nkl = 0;
        if (kk < it.lcount) {
            tl = item_tasks[it.lfirst + kk];
            if (!flags || ((accept >> flags[tl]) & 1)) nkl = tasks[tl].nket;
        }
        const int cnt = min(64, (it.lcount - base + 3) / 4);
        int j = 0;
        for (; j + 8 <= cnt; j += 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int t = __builtin_amdgcn_readlane(tl, j + u);
                const int nket = __builtin_amdgcn_readlane(nkl, j + u);
                v[u] = lane < nket ? Je[(long long)(t - it.tfirst) * stride] : 0.0;
                any |= lane < nket;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) acc += v[u];
        }
        for (; j < cnt; j++) {
            const int t = __builtin_amdgcn_readlane(tl, j);
            const int nket = __builtin_amdgcn_readlane(nkl, j);
            if (lane < nket) {
                acc += Je[(long long)(t - it.tfirst) * stride];
                any = 1;
            }
        }
    }
    part[wave][lane] = acc;
    seen[wave][lane] = any;
    __syncthreads();
    if (wave != 0) return;
    acc = part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
    any = seen[0][lane] | seen[1][lane] | seen[2][lane] | seen[3][lane];
    if (!any) return;
    const int4 pi = B.pair_info[it.ket0 + lane];
    const int fc = e / it.nfd, fd = e - fc * it.nfd;
    atomicAdd(&M[(size_t)(B.ao_off[pi.x] + fc) * nao + B.ao_off[pi.y] + fd], acc);
}

hipError_t launch_jreduce(const JRedItem *items, int nitem, int max_nfcd, const int *item_tasks, const JKTask *tasks,
                          const unsigned char *flags, int accept, const double *J, const DevBasis &B, double *M,
                          int nao, hipStream_t st)
{
    if (nitem <= 0) return hipSuccess;
    if (max_nfcd < 1 || max_nfcd > 36) return hipErrorInvalidValue;
    hipLaunchKernelGGL(jreduce_kernel, dim3(nitem, max_nfcd), dim3(256), 0, st, items, nitem, item_tasks, tasks, flags,
                       accept, J, B, M, nao);
    return hipGetLastError();
}

hipError_t launch_symmetrize(const double *M, double *G, int nao, hipStream_t st)
{
    const size_t nn = (size_t)nao * nao;
    hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, M, G, nao);
    return hipGetLastError();
}

} // namespace librint
