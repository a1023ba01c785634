// Device-side data layout shared by the host context and the kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace librint {

struct PrimPair;

// Basis data resident in HBM for the lifetime of a context.
//   pair_info[p] = {shell a, shell b, first primitive pair, primitive-pair count}
//   pair_rb[p]   = {Rb, Schwarz bound Q_p}   (Rb = VRR centre)
//   pair_rab[p]  = {Rb - Ra, 0}
//   pp[...]      = primitive-pair records (eri_core.h)
//   ao_off[s]    = first cartesian AO of shell s
//   pair_ao[p]   = {ao_off[shell a], ao_off[shell b]}
//   pp0[p]       = the first primitive-pair record of pair p (a copy of
//                  pp[pair_info[p].z]), so that a kernel's first primitive
//                  load does not wait on the pair's record offset
struct DevBasis {
    const int4 *pair_info;
    const double4 *pair_rb;
    const double4 *pair_rab;
    const PrimPair *pp;
    const int *ao_off;
    const int2 *pair_ao;
    const PrimPair *pp0;
};

// One shell pair as the kernels read it: shells, primitive-pair records
// [off, off + npp), AO offsets of the two shells, Rb and Rb - Ra.
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
    const int2 ao = B.pair_ao[p];
    v.aoa = ao.x;
    v.aob = ao.y;
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

// One wavefront of the Fock kernel: bra pair x kets [ket0, ket0 + nket) of
// the pair list (all kets of one bin).  prims = 0: every primitive pair of
// the bra; else the bra's primitive pairs [prims & 0xffff, prims >> 16) only
// (a heavy quartet's primitive sum split over several tasks, whose J/K
// contributions add up).
struct JKTask {
    int bra;
    int ket0;
    int nket;
    int prims;
};

// Bra primitive-pair range [b0, b1) of a task whose bra has npp of them.
__host__ __device__ inline void task_prims(const JKTask &t, int npp, int &b0, int &b1)
{
    if (t.prims) {
        b0 = t.prims & 0xffff;
        b1 = t.prims >> 16;
    } else {
        b0 = 0;
        b1 = npp;
    }
}

// All tasks of one bra pair within one class-pair launch, [first, first+count).
struct JKGroup {
    int bra;
    int first;
    int count;
    int pad;
};

// One thread of the list kernel.
struct ListTask {
    int bra;
    int ket;
    long long out;
};

// Per-task precision choice of the mixed Fock kernels (kernels.hip,
// launch_mix_bounds): a task runs FP32 recurrences when
// Q_b max(Dpair[b] Q[ket0], Wqd[ket0]) < tau_mix, an upper bound of every
// quartet's Q_ab Q_cd max(|D| over its six shell blocks); prec[t] receives
// 0 (FP64) or 1 (FP32).
struct MixArgs {
    const double *Dpair;
    const double *Wqd;
    double tau_mix;
    unsigned char *prec;
};

// Ket-side Coulomb slabs: instead of one global atomic per quartet into
// M[c][d], a J/K launch stores each quartet's ket-pair Coulomb block in the
// slab of its task, J[jbase + (t - tfirst)*64*NFCD + e*64 + lane]
// (e = fc*NFD + fd, lane = ket - ket0), and the reduction sums the slabs of
// every task that shares a 64-ket window (same ket0) before one atomic per
// element.  One item per (launch, window).
struct JRedItem {
    long long jbase;  // slab offset of the launch
    int tfirst;       // first task of the launch (task index space of item_tasks)
    int ket0;         // first ket pair of the window
    int nfc, nfd;     // ket shell sizes
    int lfirst, lcount;  // range in item_tasks
};

bool class_supported(int la, int lb, int lc, int ld);
// Task flags: 0 FP64, 1 FP32 recurrences, 2 skipped; a launch processes the
// tasks whose flag bit is set in `accept` (flags == nullptr: all tasks).
// counter: one int per launch, zero before the launch (launch_zero_counters)
// jslab: slab base of this launch (null: ket Coulomb block by global atomics);
// slab of task t = jslab + (t - tfirst)*64*NFC*NFD with t indexing `tasks`.
// The flat kernel is FP64 only (fp32 is ignored).
hipError_t launch_jk(int la, int lb, int lc, int ld, int fp32, const JKTask *tasks, int ntask, int *counter,
                     const unsigned char *flags, int accept, const DevBasis &B, const double *D, double *M, int nao,
                     double *jslab, int tfirst, hipStream_t st);
hipError_t launch_zero_counters(int *p, int n, hipStream_t st);
hipError_t launch_zero_doubles(double *p, size_t n, hipStream_t st);
bool class_has_fp32(int la, int lb, int lc, int ld);
// classes whose contracted block the group kernel evaluates in chunks; they
// fall back to the team kernel (FP64, no Coulomb slabs) where the group
// kernel's LDS rows do not fit
bool class_split(int la, int lb, int lc, int ld);
// grouped form (exchange rows of the bra in LDS); jk_group_lds gives its LDS
// bytes per workgroup, (size_t)-1 for classes without a grouped kernel
// (drow: with the bra's density rows staged in LDS as well)
// mixed: the form that picks FP32 recurrences per task (MixArgs)
size_t jk_group_lds(int la, int lb, int lc, int ld, int mixed, int nao, bool drow);
// wavefronts per workgroup of the grouped kernel of a class (with or without
// the density rows in LDS)
int jk_group_waves(int la, int lb, int lc, int ld, bool drow);
hipError_t launch_jk_group(int la, int lb, int lc, int ld, int mixed, const JKGroup *groups, int ngroup,
                           const JKTask *tasks, int *counter, const unsigned char *flags, const unsigned char *gmask,
                           int accept, const DevBasis &B, const double *D, double *M, int nao, double *jslab,
                           int tfirst, bool drow, const MixArgs &mx, hipStream_t st);
// per-build bounds of the mixed kernels from Dsh (Drow: nsh, Dpair, qd, Wqd: npair)
hipError_t launch_mix_bounds(const double *Dsh, int nsh, const DevBasis &B, int npair, double *Drow, double *Dpair,
                             double *qd, double *Wqd, hipStream_t st);
// shell-pair block maxima of |D| (Dsh[i*nsh + j], virtual segmented shells)
hipError_t launch_dsh(const double *D, const int *ao_off, int nsh, int nao, double *Dsh, hipStream_t st);
// sum the Coulomb slabs of every item's tasks into M (tasks: global task array)
hipError_t launch_jreduce(const JRedItem *items, int nitem, int max_nfcd, const int *item_tasks, const JKTask *tasks,
                          const unsigned char *flags, int accept, const double *J, const DevBasis &B, double *M,
                          int nao, hipStream_t st);
// per-build task classification from the shell-block maxima Dsh (launch_dsh
// first): flags (0 FP64, 1 below tau_mix, 2 below tau_skip) and group masks
hipError_t launch_classify(const JKTask *tasks, int ntask, const int *task_group, int ngroup, const DevBasis &B,
                           const double *D, const int *ao_off, int nsh, int nao, double *Dsh, double tau_mix,
                           double tau_skip, unsigned char *flags, unsigned int *gmask_w, unsigned char *gmask,
                           hipStream_t st);
hipError_t launch_list(int la, int lb, int lc, int ld, const ListTask *tasks, int ntask, const DevBasis &B,
                       double *out, int mode, int nao, hipStream_t st);
hipError_t launch_symmetrize(const double *M, double *G, int nao, hipStream_t st);

} // namespace librint
