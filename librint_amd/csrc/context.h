// Persistent two-electron context: basis staged in HBM once, Schwarz bounds
// and the binned quartet task lists computed once, then one fused J/K build
// per SCF iteration.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "eri_core.h"
#include "hl.h"
#include "kernels.h"

namespace librint {

struct Shell {
    int atom, l, nprim, nctr, pexp, pcoef;
    double r[3];
    int ao;  // first cartesian AO
    int nf;  // cartesian functions per contraction
};

struct Stats {
    long long quartets = 0;        // canonical shell quartets evaluated per build
    long long prim_quartets = 0;   // primitive quartets evaluated per build
    long long integrals = 0;       // contracted cartesian integrals per build
    double flops = 0;              // algorithmic flops per build (fp64 path)
    long long candidates = 0;      // canonical shell quartets before screening
    int launches = 0;
    float last_kernel_ms = 0.f;    // device time of the last build (events)
    float last_grad_ms = 0.f;      // device time of the last adjoint (grad2e) launches
    int scf_purified = 0;          // densities by purification in the last SCF on the context
    int streams = 0;               // streams of the last build's class launches
};

struct ClassLaunch {
    int la, lb, lc, ld;
    int first, count;    // range in the device task array
    int gfirst, gcount;  // range in the device group array
    int fp32 = 0;        // mixed: FP32 recurrences for the tasks below tau_mix, decided per build
    long long quartets = 0, prim_quartets = 0, lanes = 0;
    double flops = 0;
    float ms = 0.f;    // device time of this launch in the last build
    float reduce_ms = 0.f;  // ... of which its slab chunks' Coulomb reductions
    int variant = -1;  // kernel of the last build: 0 jk_kernel, 1 jk_group_kernel, 2 jk_team_kernel, 3 hl_kernel
    long long jbase = -1;  // Coulomb slab offset in the whole-build layout (-1: global atomics)
    int cfirst = 0, ccount = 0;  // range in the slab chunks (one chunk when the slabs are kept whole)
    int drow = 0;                // bra density rows staged in LDS in the last build
    int drow_plan = 0;           // ... as planned with the task lists (group sizes follow it)
};

class Context {
public:
    Context(const int *atm, size_t atm_l, const int *bas, size_t bas_l, const double *env, size_t env_l);
    ~Context();

    int nao() const { return nao_; }
    int nbas() const { return (int)shells_.size(); }
    int natm() const { return (int)(atm_.size() / 6); }

    // Screening threshold on Q_ij * Q_kl; rebuilds the task lists.
    void set_schwarz(double tau);
    double schwarz() const { return tau_; }
    // One device's share of the Fock build when nshards devices split it:
    // the groups of every class launch are dealt, heaviest first, each to
    // the share with the least primitive-quartet cost so far, and this
    // context keeps share `shard` (0 <= shard < nshards).  fock_* then
    // return that share's G; the G's of all shares sum to the full build.
    // nshards = 1 (default) is the whole build.  Rebuilds the task lists.
    void set_shard(int nshards, int shard);
    int nshards() const { return nshards_; }
    // Precision mode 0: FP64 everywhere.  Mode 1 (mixed): every build bounds
    // each 64-ket task by the largest Q_ab Q_cd max(|D| over the six blocks
    // the quartet reads) of its kets, from that build's density; tasks below
    // tau_mix run FP32 recurrences.  tau_skip > 0 drops, per build, the tasks
    // below it (density screening, both modes).
    void set_precision(int mode, double tau_mix, double tau_skip);
    int precision() const { return precision_; }
    // Quartets of the last build by task flag (0 FP64, 1 FP32, 2 skipped).
    void flag_counts(long long *quartets3, long long *prim3);

    // G = J - K/2 for a symmetric density (device pointers, nao*nao, row-major).
    void fock_device(const double *D_dev, double *G_dev, hipStream_t st);
    // The same build using `len` doubles of caller-provided device scratch
    // (the accumulator M, the shell-block density maxima, the work counters
    // and the f/g classes' block scratch) instead of the context's buffers:
    // nothing is allocated, one stream, no Coulomb slabs.  Throws when the
    // scratch is smaller than scratch_doubles_needed().
    void fock_device_scratch(const double *D_dev, double *G_dev, hipStream_t st, double *scratch, size_t len);
    size_t scratch_doubles_needed() const;
    // Host-pointer convenience wrapper.
    void fock_host(const double *D, double *G);

    // Full nao^4 tensor (host buffer, R[((i*n+j)*n+k)*n+l]).
    void tensor_host(double *R);
    // Same into a device buffer of nao^4 doubles.
    void tensor_device(double *dR);
    // Blocks for a list of shell quartets, each in cint2e_cart layout
    // (i fastest), concatenated; returns the number of doubles written.
    size_t shells_host(const int *quartets, int nq, double *out);
    // Schwarz bound per shell pair (i >= j), host copy, index i*(i+1)/2 + j.
    std::vector<double> schwarz_pairs() const;
    // Adjoint of E2 = 1/2 tr(P G[P]) over the screened quartet lists (FP64):
    // gatom[3*natm] = dE2/dR, denv[env_l] = dE2/d(env entry) for every
    // primitive exponent and contraction coefficient (zero elsewhere).
    void grad2e(const double *P, double *gatom, double *denv, size_t env_l);

    const Stats &stats() const { return stats_; }
    void note_scf_purified(int k) { stats_.scf_purified = k; }
    const std::vector<Shell> &shells() const { return shells_; }
    const std::vector<int> &atm() const { return atm_; }
    const std::vector<double> &env() const { return env_; }
    const std::vector<int> &bas() const { return bas_; }
    const std::vector<ClassLaunch> &launches() const { return launches_; }
    // per-launch device times of the last build (host wrapper synchronises)
    void collect_class_times();

private:
    void build_pairs();
    void compute_schwarz();
    void build_tasks();
    int default_streams() const;
    bool lane_path(int la, int lb, int lc, int ld, int mixed) const;
    bool drow_plan(int la, int lb, int lc, int ld, int fp32, double prim_quartets, double quartets) const;
    void upload();
    void ensure_side_streams(int n);
    struct FockScratch {
        double *M, *dsh, *mix, *hl;
        int *counters;
        size_t hl_len;
    };
    void fock_impl(const double *D_dev, double *G_dev, hipStream_t st, const FockScratch *sc);

    std::vector<int> atm_, bas_;
    std::vector<double> env_;
    std::vector<Shell> shells_;
    int nao_ = 0;
    double tau_ = 1e-14;
    int nshards_ = 1, shard_ = 0;

    // pair list in bin order (bins = (class, npp), Q descending inside a bin)
    struct PairH {
        int sa, sb;   // sa: HRR target (la <= lb), sb: VRR centre
        int shell_i, shell_j;  // canonical i >= j
        int cls, npp, off;
        double Q;
    };
    std::vector<PairH> pairs_;
    std::vector<PrimPair> pp_;
    std::vector<int> pp_idx_;      // primitive indices (on sa, on sb) per record
    std::vector<int> pair_of_ij_;  // canonical (i,j) -> position in pairs_
    struct Bin {
        int cls, npp, first, count;
    };
    std::vector<Bin> bins_;

    std::vector<ClassLaunch> launches_;
    // consecutive groups of one launch whose Coulomb slabs fit the slab
    // budget, with their reduction items; jbase -1: no slabs (atomics)
    struct SlabChunk {
        int gfirst, gcount, tfirst, tcount, ifirst, icount;
        long long jbase;
    };
    std::vector<SlabChunk> chunks_;
    Stats stats_;

    // device
    int4 *d_pair_info_ = nullptr;
    double4 *d_pair_rb_ = nullptr, *d_pair_rab_ = nullptr;
    PrimPair *d_pp_ = nullptr;
    int *d_ao_off_ = nullptr;
    int2 *d_pair_ao_ = nullptr;    // AO offsets of each pair's two shells
    PrimPair *d_pp0_ = nullptr;    // first primitive-pair record of each pair
    JKTask *d_tasks_ = nullptr;
    JKGroup *d_groups_ = nullptr;
    int *d_task_group_ = nullptr;        // group of each task
    unsigned char *d_flags_ = nullptr;   // per task, per build
    unsigned int *d_gmask_w_ = nullptr;  // per group, per build
    unsigned char *d_gmask_ = nullptr;
    double *d_dsh_ = nullptr;            // shell-block max |D|, per build
    double *d_mixb_ = nullptr;           // mixed-precision bounds per build: Drow, Dpair, Q Dpair, Wqd
    int precision_ = 0;
    double tau_mix_ = 0.0, tau_skip_ = 0.0;
    bool classified_ = false;            // flags valid for the last build
    bool mixed_last_ = false;            // the last build ran the mixed kernels (flags hold 0/1)
    std::vector<JKTask> h_tasks_;        // host copy (counts by flag)
    int ngroups_ = 0;
    int *d_counters_ = nullptr;  // one work counter per class launch
    int max_lds_ = 0;
    bool no_group_ = false;      // LIBRINT_NO_GROUP: flat per-task kernels only
    double *d_M_ = nullptr, *d_D_ = nullptr, *d_G_ = nullptr;
    // ket-side Coulomb slabs and their reduction items (kernels.h)
    double *d_J_ = nullptr;
    size_t nJ_ = 0;
    JRedItem *d_items_ = nullptr;
    int *d_item_tasks_ = nullptr;
    int nitems_ = 0, max_nfcd_ = 1;
    bool jslab_whole_ = true;     // one reduction after all launches (else one per launch)
    bool no_jslab_ = false;       // LIBRINT_NO_JSLAB: global atomics only
    int drow_mode_ = 2;           // bra density rows in LDS: 0 never, 1 where they fit, 2 cheap quartets (LIBRINT_DROW)
    int2 *d_pp_idx_ = nullptr;      // adjoint kernel: primitive indices per record
    int2 *d_shell_slot_ = nullptr;  // {atom, first primitive slot} per shell
    double *d_slot_coef_ = nullptr; // contraction coefficient per slot
    int nslot_ = 0;
    // adjoint kernel, ket side: the same quartets listed ket-major (the task's
    // fixed pair is the canonical ket, its window runs over canonical bras)
    struct GradLaunch {
        int la, lb, lc, ld, first, count;
    };
    std::vector<GradLaunch> glaunch1_;
    JKTask *d_gtasks1_ = nullptr;
    // bra side of the chunked classes: their Fock tasks (64-ket windows) cut
    // into grad_kets_ windows, since the adjoint kernel takes a task's kets
    // one at a time (count 0: the class uses its Fock tasks as they are)
    std::vector<GradLaunch> glaunch0_;
    JKTask *d_gtasks0_ = nullptr;
    bool gtasks_valid_ = false;
    void build_grad_tasks();
    // classes with an f or g shell (hl.hip): scratch of one contracted block
    // per workgroup, allocated on first use
    double *d_hlscr_ = nullptr;
    size_t nhlscr_ = 0;
    double *hl_scratch(const HLShape &S);
    void hl_scratch_for(const std::vector<std::array<int, 4>> &classes);  // size for the largest
    hipError_t list_launch(int la, int lb, int lc, int ld, const ListTask *tasks, int ntask, const DevBasis &B,
                           double *out, int mode, hipStream_t st);
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    std::vector<hipEvent_t> evc_;  // one event after each class launch
    std::vector<hipEvent_t> evs_;  // one event before each class launch (same stream)
    // around each slab chunk's Coulomb reduction (inside its launch's events)
    std::vector<hipEvent_t> evr0_, evr1_;
    std::vector<char> red_rec_;    // chunk's reduction recorded in the last build
    // class launches after the dominant one are dealt round-robin over the
    // caller's stream and nstreams_ - 1 extra streams (LIBRINT_STREAMS),
    // joined before the Coulomb reduction: small launches and launch tails
    // overlap
    int nstreams_ = 0;      // 0: default_streams()
    // ket window of the Fock tasks of the team and f/g classes
    // (LIBRINT_TEAM_KETS; measured on (H2O)32 with three streams: 4 gives
    // 35.39-35.58 ms, 8 35.75-35.85, 2 35.91, 16 35.62)
    int team_kets_ = 4;
    // ket window of the adjoint's short tasks: the ket-major lists of those
    // classes and the chunked classes' Fock tasks cut for the bra side
    // (LIBRINT_GRAD_KETS; 4 instead of 8 measured 19.95-20.11 against
    // 19.82-19.89 ms per (H2O)8 gradient)
    int grad_kets_ = 8;
    // groups per launch the group size leaves at least (LIBRINT_GROUP_MIN):
    // small groups balance a launch, large ones stage and flush their rows
    // less often; measured with three streams and the launch order by flops
    // per task, (H2O)32 32.8-32.9 ms at 256, 33.0 at 512, 33.0 at 1024,
    // 33.5 at 2048 and 35.1 at 4096; C16H34 1.455-1.459 ms at 256,
    // 1.457-1.458 at 512 and 1.483-1.487 at 4096
    int group_min_ = 256;
    // lane-per-quartet classes: a task whose lanes would each evaluate more
    // than prim_split_ primitive quartets is split by bra primitive pairs
    // (LIBRINT_PRIM_SPLIT, default 48; 0 = never), so that the heaviest
    // quartets (up to 36 x 36 primitive pairs on one lane) do not end their
    // launch as one wavefront's long tail
    int prim_split_ = 48;
    int dominant_ = 0;      // launch with the most flops: runs first, alone
    std::vector<int> launch_order_;  // dominant_, then the others by decreasing flops per task
    int slab_streams_ = 1;  // per-launch slab mode: buffers of jslot_ doubles, one per stream
    size_t jslot_ = 0;
    std::vector<hipStream_t> side_;
    std::vector<hipEvent_t> evj_;  // join events, one per extra stream
    hipEvent_t evf_ = nullptr;     // fork event
    DevBasis devb() const;
};

int cls_of(int la, int lb);  // pair class index, la <= lb
void check(hipError_t e, const char *what);

}  // namespace librint
