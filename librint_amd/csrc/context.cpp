// Host side of the two-electron context: basis parsing (reference
// src/utils.rs:17-81 layout, src/eri.rs:98-227 per-quartet constants),
// primitive-pair staging, Schwarz bounds from the GPU diagonal (ij|ij),
// binning and quartet task lists.
#include "context.h"

#include "grad.h"
#include "hl.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <tuple>

namespace librint {

void check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// pair classes (la <= lb), lb major: (00) (01) (11) (02) (12) (22) (03) ...
// (33) (04) ... (44); classes 0-5 have their own kernels, 6-14 (an f or g
// shell) use the run-time shaped kernels of hl.hip
int cls_of(int la, int lb)
{
    if (la < 0 || la > lb || lb > HL_LMAX) return -1;
    return lb * (lb + 1) / 2 + la;
}

static void cls_ls(int c, int &la, int &lb)
{
    lb = 0;
    while ((lb + 1) * (lb + 2) / 2 <= c) lb++;
    la = c - lb * (lb + 1) / 2;
}

static constexpr int NCLS = (HL_LMAX + 1) * (HL_LMAX + 2) / 2;

// angular classes (la, lb, lc, ld) of a (bra class, ket class) keyed list
template <typename M>
static std::vector<std::array<int, 4>> list_classes(const M &by_cls)
{
    std::vector<std::array<int, 4>> out;
    for (const auto &kv : by_cls) {
        std::array<int, 4> c;
        cls_ls(kv.first.first, c[0], c[1]);
        cls_ls(kv.first.second, c[2], c[3]);
        out.push_back(c);
    }
    return out;
}

static double common_fac_sp(int l)
{ // reference src/g1e.rs:776-782
    if (l == 0) return 0.282094791773878143;
    if (l == 1) return 0.488602511902919921;
    return 1.0;
}

template <typename T>
static T *dev_upload(const std::vector<T> &v)
{
    T *p = nullptr;
    if (v.empty()) return nullptr;
    check(hipMalloc(&p, sizeof(T) * v.size()), "hipMalloc");
    check(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return p;
}

template <typename T>
static void dev_free(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

Context::Context(const int *atm, size_t atm_l, const int *bas, size_t bas_l, const double *env, size_t env_l)
    : atm_(atm, atm + atm_l), bas_(bas, bas + bas_l), env_(env, env + env_l)
{
    if (atm_l % 6 || bas_l % 8) throw std::invalid_argument("atm/bas lengths must be multiples of 6/8");
    if (env_l > 8 && env_[8] != 0.0) throw std::invalid_argument("range-separated (env[8] != 0) integrals are not supported");
    const int nb = (int)(bas_l / 8);
    int ao = 0;
    for (int s = 0; s < nb; s++) {
        const int *b = &bas_[8 * s];
        if (b[0] < 0 || (size_t)b[0] >= atm_l / 6) throw std::invalid_argument("bas atom index out of range");
        if (b[1] < 0 || b[1] > HL_LMAX) throw std::invalid_argument("angular momentum l > 4 is not supported by the GPU path");
        // general contraction: one segmented virtual shell per contracted function
        for (int c = 0; c < b[3]; c++) {
            Shell sh;
            sh.atom = b[0];
            sh.l = b[1];
            sh.nprim = b[2];
            sh.nctr = 1;
            sh.pexp = b[5];
            sh.pcoef = b[6] + c * b[2];
            const int pr = atm_[6 * b[0] + 1];
            for (int d = 0; d < 3; d++) sh.r[d] = env_[pr + d];
            sh.nf = ncart(b[1]);
            sh.ao = ao;
            ao += sh.nf;
            shells_.push_back(sh);
        }
    }
    nao_ = ao;
    no_group_ = getenv("LIBRINT_NO_GROUP") != nullptr;
    no_jslab_ = getenv("LIBRINT_NO_JSLAB") != nullptr;
    if (const char *e = getenv("LIBRINT_DROW")) drow_mode_ = std::max(0, std::min(2, atoi(e)));
    check(hipEventCreate(&ev0_), "hipEventCreate");
    check(hipEventCreate(&ev1_), "hipEventCreate");
    if (const char *e = getenv("LIBRINT_STREAMS")) nstreams_ = std::max(1, std::min(8, atoi(e)));
    if (const char *e = getenv("LIBRINT_TEAM_KETS")) team_kets_ = std::max(1, std::min(64, atoi(e)));
    if (const char *e = getenv("LIBRINT_GRAD_KETS")) grad_kets_ = std::max(1, std::min(64, atoi(e)));
    if (const char *e = getenv("LIBRINT_GROUP_MIN")) group_min_ = std::max(1, atoi(e));
    if (const char *e = getenv("LIBRINT_PRIM_SPLIT")) prim_split_ = std::max(0, atoi(e));
    {
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&max_lds_, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "hipDeviceGetAttribute");
    }
    build_pairs();
    upload();
    compute_schwarz();
    build_tasks();
}

Context::~Context()
{
    dev_free(d_pair_info_);
    dev_free(d_pair_rb_);
    dev_free(d_pair_rab_);
    dev_free(d_pair_ao_);
    dev_free(d_pp0_);
    dev_free(d_pp_);
    dev_free(d_ao_off_);
    dev_free(d_tasks_);
    dev_free(d_groups_);
    dev_free(d_task_group_);
    dev_free(d_flags_);
    dev_free(d_gmask_w_);
    dev_free(d_gmask_);
    dev_free(d_dsh_);
    dev_free(d_mixb_);
    dev_free(d_counters_);
    dev_free(d_M_);
    dev_free(d_D_);
    dev_free(d_G_);
    dev_free(d_J_);
    dev_free(d_items_);
    dev_free(d_item_tasks_);
    dev_free(d_pp_idx_);
    dev_free(d_shell_slot_);
    dev_free(d_slot_coef_);
    dev_free(d_gtasks1_);
    dev_free(d_gtasks0_);
    dev_free(d_hlscr_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    for (hipEvent_t e : evc_) (void)hipEventDestroy(e);
    for (hipEvent_t e : evs_) (void)hipEventDestroy(e);
    for (hipEvent_t e : evr0_) (void)hipEventDestroy(e);
    for (hipEvent_t e : evr1_) (void)hipEventDestroy(e);
    for (hipEvent_t e : evj_) (void)hipEventDestroy(e);
    if (evf_) (void)hipEventDestroy(evf_);
    for (hipStream_t q : side_) (void)hipStreamDestroy(q);
}

DevBasis Context::devb() const
{
    DevBasis B;
    B.pair_info = d_pair_info_;
    B.pair_rb = d_pair_rb_;
    B.pair_rab = d_pair_rab_;
    B.pp = d_pp_;
    B.ao_off = d_ao_off_;
    B.pair_ao = d_pair_ao_;
    B.pp0 = d_pp0_;
    return B;
}

void Context::build_pairs()
{
    const double expcutoff = (env_[0] == 0.0) ? 60.0 : std::max(40.0, env_[0]) + 1.0;
    const int ns = (int)shells_.size();
    pairs_.clear();
    pp_.clear();
    pp_idx_.clear();
    pair_of_ij_.assign((size_t)ns * (ns + 1) / 2, -1);
    for (int i = 0; i < ns; i++)
        for (int j = 0; j <= i; j++) {
            const Shell &si = shells_[i], &sj = shells_[j];
            PairH P;
            P.shell_i = i;
            P.shell_j = j;
            if (si.l <= sj.l) {
                P.sa = i;
                P.sb = j;
            } else {
                P.sa = j;
                P.sb = i;
            }
            const Shell &A = shells_[P.sa], &Bs = shells_[P.sb];
            P.cls = cls_of(A.l, Bs.l);
            double rab2 = 0;
            for (int d = 0; d < 3; d++) rab2 += (A.r[d] - Bs.r[d]) * (A.r[d] - Bs.r[d]);
            // pair-level part of the primitive screen (reference src/eri.rs:192-203)
            const double alast = env_[A.pexp + A.nprim - 1] + env_[Bs.pexp + Bs.nprim - 1];
            double log_rr = 1.7 - 1.5 * std::log(alast);
            if (A.l + Bs.l > 0) log_rr += (A.l + Bs.l) * std::log(std::sqrt(rab2) + 1.0);
            const double cf = common_fac_sp(A.l) * common_fac_sp(Bs.l);
            P.off = (int)pp_.size();
            for (int q = 0; q < Bs.nprim; q++)
                for (int p = 0; p < A.nprim; p++) {
                    const double ap = env_[A.pexp + p], aq = env_[Bs.pexp + q];
                    const double cp = env_[A.pcoef + p], cq = env_[Bs.pcoef + q];
                    if (cp == 0.0 || cq == 0.0) continue;
                    const double e = ap + aq;
                    const double eij = ap * aq / e * rab2;
                    const double cce = eij - log_rr - std::log(std::fabs(cp)) - std::log(std::fabs(cq));
                    if (cce > expcutoff) continue;
                    PrimPair r;
                    r.e = e;
                    // sqrt(2 pi^{5/2}): the quartet prefactor splits evenly over the two pairs
                    r.K = 5.914967172795612 * cp * cq * cf * std::exp(-eij) / e;
                    r.h = 0.5 / e;
                    r.pbx = (ap * A.r[0] + aq * Bs.r[0]) / e - Bs.r[0];
                    r.pby = (ap * A.r[1] + aq * Bs.r[1]) / e - Bs.r[1];
                    r.pbz = (ap * A.r[2] + aq * Bs.r[2]) / e - Bs.r[2];
                    r.pad0 = ap;  // exponents, read by the adjoint kernel
                    r.pad1 = aq;
                    pp_.push_back(r);
                    pp_idx_.push_back(p);
                    pp_idx_.push_back(q);
                }
            P.npp = (int)pp_.size() - P.off;
            P.Q = 0.0;
            if (P.npp == 0) continue;
            pair_of_ij_[(size_t)i * (i + 1) / 2 + j] = (int)pairs_.size();
            pairs_.push_back(P);
        }
}

void Context::upload()
{
    dev_free(d_pair_info_);
    dev_free(d_pair_rb_);
    dev_free(d_pair_rab_);
    dev_free(d_ao_off_);
    std::vector<int4> info(pairs_.size());
    std::vector<double4> rb(pairs_.size()), rab(pairs_.size());
    for (size_t p = 0; p < pairs_.size(); p++) {
        const PairH &P = pairs_[p];
        const Shell &A = shells_[P.sa], &Bs = shells_[P.sb];
        info[p] = make_int4(P.sa, P.sb, P.off, P.npp);
        rb[p] = make_double4(Bs.r[0], Bs.r[1], Bs.r[2], P.Q);
        rab[p] = make_double4(Bs.r[0] - A.r[0], Bs.r[1] - A.r[1], Bs.r[2] - A.r[2], 0.0);
    }
    std::vector<int> aoo(shells_.size() + 1);
    for (size_t s = 0; s < shells_.size(); s++) aoo[s] = shells_[s].ao;
    aoo[shells_.size()] = nao_;
    std::vector<int2> pao(pairs_.size());
    std::vector<PrimPair> pp0(pairs_.size());
    for (size_t p = 0; p < pairs_.size(); p++) {
        pao[p] = make_int2(shells_[pairs_[p].sa].ao, shells_[pairs_[p].sb].ao);
        pp0[p] = pp_[pairs_[p].off];
    }
    dev_free(d_pair_ao_);
    dev_free(d_pp0_);
    d_pair_ao_ = dev_upload(pao);
    d_pp0_ = dev_upload(pp0);
    d_pair_info_ = dev_upload(info);
    d_pair_rb_ = dev_upload(rb);
    d_pair_rab_ = dev_upload(rab);
    d_ao_off_ = dev_upload(aoo);
    if (!d_pp_) d_pp_ = dev_upload(pp_);
}

void Context::compute_schwarz()
{
    // diagonal (ij|ij) per pair on the GPU, grouped by class
    const int np = (int)pairs_.size();
    std::vector<ListTask> tasks;
    std::vector<std::pair<int, int>> ranges(NCLS, {0, 0});
    for (int c = 0; c < NCLS; c++) {
        ranges[c].first = (int)tasks.size();
        for (int p = 0; p < np; p++)
            if (pairs_[p].cls == c) tasks.push_back(ListTask{p, p, (long long)p});
        ranges[c].second = (int)tasks.size() - ranges[c].first;
    }
    ListTask *dt = dev_upload(tasks);
    double *dq = nullptr;
    check(hipMalloc(&dq, sizeof(double) * std::max(np, 1)), "hipMalloc");
    const DevBasis B = devb();
    {
        std::vector<std::array<int, 4>> cl;
        for (int c = 0; c < NCLS; c++)
            if (ranges[c].second) {
                int la, lb;
                cls_ls(c, la, lb);
                cl.push_back({la, lb, la, lb});
            }
        hl_scratch_for(cl);
    }
    for (int c = 0; c < NCLS; c++) {
        if (!ranges[c].second) continue;
        int la, lb;
        cls_ls(c, la, lb);
        check(list_launch(la, lb, la, lb, dt + ranges[c].first, ranges[c].second, B, dq, 1, 0), "schwarz launch");
    }
    std::vector<double> Q(np);
    check(hipMemcpy(Q.data(), dq, sizeof(double) * np, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    dev_free(dq);
    dev_free(dt);
    for (int p = 0; p < np; p++) pairs_[p].Q = Q[p];

    // reorder into bins: (class, npp) ascending, Q descending inside a bin
    std::vector<int> order(np);
    for (int p = 0; p < np; p++) order[p] = p;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const PairH &a = pairs_[x], &b = pairs_[y];
        if (a.cls != b.cls) return a.cls < b.cls;
        if (a.npp != b.npp) return a.npp < b.npp;
        return a.Q > b.Q;
    });
    std::vector<PairH> sorted(np);
    for (int p = 0; p < np; p++) sorted[p] = pairs_[order[p]];
    pairs_.swap(sorted);
    std::fill(pair_of_ij_.begin(), pair_of_ij_.end(), -1);
    bins_.clear();
    for (int p = 0; p < np; p++) {
        const PairH &P = pairs_[p];
        pair_of_ij_[(size_t)P.shell_i * (P.shell_i + 1) / 2 + P.shell_j] = p;
        if (bins_.empty() || bins_.back().cls != P.cls || bins_.back().npp != P.npp)
            bins_.push_back(Bin{P.cls, P.npp, p, 0});
        bins_.back().count++;
    }
    upload();
}

double *Context::hl_scratch(const HLShape &S)
{
    // one contracted block per workgroup; up to 8192 workgroups, at most
    // 2^26 doubles (512 MiB).  Callers size it for their largest class before
    // queueing any launch: a reallocation here would free memory that an
    // earlier launch may still be using.
    const size_t want = std::min<size_t>((size_t)1 << 26, (size_t)S.nf * 8192);
    if (want > nhlscr_) {
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        dev_free(d_hlscr_);
        check(hipMalloc(&d_hlscr_, sizeof(double) * want), "hipMalloc");
        nhlscr_ = want;
    }
    return d_hlscr_;
}

void Context::hl_scratch_for(const std::vector<std::array<int, 4>> &classes)
{
    const HLShape *big = nullptr;
    HLShape best{};
    for (const auto &c : classes)
        if (hl_class(c[0], c[1], c[2], c[3])) {
            const HLShape S = hl_shape(c[0], c[1], c[2], c[3]);
            if (!big || S.nf > best.nf) {
                best = S;
                big = &best;
            }
        }
    if (big) hl_scratch(best);
}

hipError_t Context::list_launch(int la, int lb, int lc, int ld, const ListTask *tasks, int ntask, const DevBasis &B,
                                double *out, int mode, hipStream_t st)
{
    if (!hl_class(la, lb, lc, ld)) return launch_list(la, lb, lc, ld, tasks, ntask, B, out, mode, nao_, st);
    const HLShape S = hl_shape(la, lb, lc, ld);
    return launch_hl_list(S, tasks, ntask, B, out, mode, nao_, hl_scratch(S), nhlscr_, st);
}

void Context::set_schwarz(double tau)
{
    tau_ = tau;
    build_tasks();
    gtasks_valid_ = false;
}

void Context::set_shard(int nshards, int shard)
{
    if (nshards < 1 || shard < 0 || shard >= nshards) throw std::invalid_argument("shard must satisfy 0 <= shard < nshards");
    nshards_ = nshards;
    shard_ = shard;
    build_tasks();
    gtasks_valid_ = false;
}

// Per primitive quartet flop count of the fp64 kernel's algorithm (the
// definition used for roofline numbers): root evaluation, recurrence
// coefficients, VRR, HRR and the gout products, FMA counted as 2 flops.
// An all-s quartet evaluates the weight polynomial only (degree 11, Estrin:
// 11 FMA) after the segment variable (6), no node and no recurrence
// coefficient, and adds w fac1 (1 mul) into its one integral (1 add).
static double flops_prim(int la, int lb, int lc, int ld)
{
    const int nmax = la + lb, mmax = lc + ld, nr = (nmax + mmax) / 2 + 1;
    if (nmax + mmax == 0) return 30 + 6 + 22 + 1 + 1;
    const int dk = la + 1, dl = dk * (lc + 1);
    (void)dl;
    const int nf = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
    double setup = 30;                 // exponents, P-Q, x, prefactor (incl. sqrt/div)
    double roots = nr * 2.0 * (11 * 2 + 2) + 4;  // 2n Clenshaw of degree 11
    double coef = 16;                  // b00 b10 b01 ta tc, c00 c0p (x3), w*fac
    double vrr = 0;
    if (nmax > 0) vrr += 1 + 4.0 * (nmax - 1);
    if (mmax > 0) {
        vrr += 1 + 4.0 * (mmax - 1);
        if (nmax > 0) vrr += 3 + 6.0 * (mmax - 1);
    }
    if (nmax > 1) vrr += (double)mmax * (nmax - 1) * 7.0;
    double hrr = 0;
    for (int i = 1; i <= la; i++) hrr += 2.0 * (nmax - i + 1) * (mmax + 1);
    for (int k = 1; k <= lc; k++) hrr += 2.0 * (lb + 1) * (mmax - k + 1) * dk;
    const double gout = 3.0 * nf;  // 2 mul + 1 add
    return setup + roots + nr * (coef + 3 * (vrr + hrr) + gout);
}

// A chunked class runs lane per quartet (64-ket tasks with Coulomb slabs,
// FP32 in mixed precision) only in the group kernel, i.e. where its LDS rows
// fit; elsewhere it takes the team kernel in FP64 with atomics.
bool Context::lane_path(int la, int lb, int lc, int ld, int mixed) const
{
    if (!class_split(la, lb, lc, ld)) return true;
    return !no_group_ && jk_group_lds(la, lb, lc, ld, mixed ? 1 : 0, nao_, false) <= (size_t)max_lds_;
}

// Staging the bra's density rows in LDS (DROW) doubles the row space per
// workgroup but turns the exchange update's gathers of D into LDS reads.
// Default (mode 2): stage them where a quartet has fewer than 4 primitive
// quartets on average (the exchange update is then a large part of the
// work) and the rows do not cost occupancy: two 4-wave workgroups fit a CU,
// or the class runs 8- or 12-wave workgroups (one of them holds the 2 or 3
// waves per SIMD its registers allow).  LIBRINT_DROW=0 never, 1 wherever it
// fits.  Measured: (H2O)32 48.5 -> 46.8 ms, the contracted C16H34 lists
// unchanged (none qualifies; staging them all was slower, 1.62 -> 1.65 ms).
bool Context::drow_plan(int la, int lb, int lc, int ld, int fp32, double prim_quartets, double quartets) const
{
    const size_t lds = jk_group_lds(la, lb, lc, ld, fp32, nao_, false);
    if (no_group_ || lds > (size_t)max_lds_) return false;  // no group kernel
    const size_t lds_drow = jk_group_lds(la, lb, lc, ld, fp32, nao_, true);
    if (lds_drow > (size_t)max_lds_) return false;
    if (drow_mode_ != 2) return drow_mode_ == 1;
    return prim_quartets < 4.0 * quartets &&
           (lds_drow <= (size_t)max_lds_ / 2 || jk_group_waves(la, lb, lc, ld, true) > 4);
}

// Streams for the class launches after the dominant one (LIBRINT_STREAMS
// overrides): measured with the launches in order of decreasing flops per
// task, three streams are best for builds of many class launches ((H2O)32,
// 21 launches: 35.79-35.83 against 35.99-36.00 ms with two; four slower)
// and two for a few (C16H34, 6 launches: 1.487-1.490 against 1.494-1.501 ms
// with three).
int Context::default_streams() const { return launches_.size() > 12 ? 3 : 2; }

void Context::build_tasks()
{
    // the lists below replace device buffers that launches of an earlier
    // build (on this context's extra streams too) may still be reading
    if (d_tasks_) check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    gtasks_valid_ = false;  // the adjoint's lists derive from these tasks
    const int nb = (int)bins_.size();
    launches_.clear();
    stats_ = Stats();
    const long long np = (long long)pairs_.size();
    stats_.candidates = np * (np + 1) / 2;
    // A group = consecutive tasks of one bra pair for one ket class; its size
    // is set per launch below (LIBRINT_GROUP_TASKS overrides).
    int GROUP_MAX_TASKS = 0;
    // (at most 256: a wavefront of the group kernel holds the flags of its
    // quarter of a group's tasks, one per lane)
    if (const char *e = getenv("LIBRINT_GROUP_TASKS")) GROUP_MAX_TASKS = std::max(1, std::min(256, atoi(e)));
    struct Acc {
        std::vector<JKTask> tasks;   // contiguous per bra pair
        std::vector<double> cost;    // primitive quartets per task
        std::vector<std::pair<double, JKGroup>> groups;
        ClassLaunch stat;
    };
    std::map<std::tuple<int, int, int>, Acc> acc;  // (bra class, ket class, 0)
    // mixed precision: the classes with FP32 kernels launch the form that
    // picks FP64 or FP32 recurrences per task from each build's density
    const bool mixed = precision_ == 1 && tau_mix_ > 0.0;
    for (int X = 0; X < nb; X++) {
        const Bin &bx = bins_[X];
        int la, lb;
        cls_ls(bx.cls, la, lb);
        for (int i = 0; i < bx.count; i++) {
            const int p = bx.first + i;
            const double Qp = pairs_[p].Q;
            for (int Y = 0; Y <= X; Y++) {
                const Bin &by = bins_[Y];
                int lc, ld;
                cls_ls(by.cls, lc, ld);
                if (!class_supported(la, lb, lc, ld) && !hl_class(la, lb, lc, ld))
                    throw std::runtime_error("angular class not supported by the GPU kernels");
                // kets of bin Y with Q_p Q_k > tau: a prefix (Q descending)
                int lo = 0, hi = by.count;
                while (lo < hi) {
                    const int mid = (lo + hi) / 2;
                    if (Qp * pairs_[by.first + mid].Q > tau_) lo = mid + 1;
                    else hi = mid;
                }
                int n = lo;
                if (X == Y) n = std::min(n, i + 1);
                if (n <= 0) continue;
                const double fp = flops_prim(la, lb, lc, ld);
                const int nf = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
                {
                    const int part = 0, k_begin = 0, k_end = n;
                    const int nn = k_end - k_begin;
                    const double fl = (double)nn * bx.npp * by.npp * fp + 12.0 * nf * nn;
                    stats_.quartets += nn;
                    stats_.prim_quartets += (long long)nn * bx.npp * by.npp;
                    stats_.integrals += (long long)nn * nf;
                    stats_.flops += fl;
                    Acc &A = acc[std::make_tuple(bx.cls, by.cls, part)];
                    A.stat.quartets += nn;
                    A.stat.prim_quartets += (long long)nn * bx.npp * by.npp;
                    A.stat.flops += fl;
                    A.stat.lanes += 64LL * ((nn + 63) / 64);
                    // lane-per-quartet classes take 64-ket windows (a lane per
                    // ket, Coulomb slabs per window); the classes whose
                    // wavefront works through its kets one at a time take
                    // short windows, so that no task runs far longer than the rest
                    const bool lane = class_has_fp32(la, lb, lc, ld) && lane_path(la, lb, lc, ld, mixed);
                    const int win = lane ? 64 : team_kets_;
                    int parts = 1;
                    if (prim_split_ > 0 && lane && bx.npp < 65536)
                        parts = std::min(bx.npp, (bx.npp * by.npp + prim_split_ - 1) / prim_split_);
                    for (int k0 = k_begin; k0 < k_end; k0 += win) {
                        const int nk = std::min(win, k_end - k0);
                        for (int q = 0; q < parts; q++) {
                            const int b0 = bx.npp * q / parts, b1 = bx.npp * (q + 1) / parts;
                            A.tasks.push_back(JKTask{p, by.first + k0, nk, parts > 1 ? (b0 | (b1 << 16)) : 0});
                            A.cost.push_back((double)nk * (b1 - b0) * by.npp);
                        }
                    }
                }
            }
        }
    }
    std::vector<JKTask> tasks;
    std::vector<JKGroup> groups;
    std::vector<int> task_group;
    for (auto &kv : acc) {
        Acc &A = kv.second;
        // Group size: a group's exchange rows ((NFA+NFB) x nao in LDS) are
        // flushed with global atomics once per group, so a group should hold
        // about (NFA+NFB)*nao/64 tasks for the flush to stay small next to its
        // quartets; at most 64 x the workgroup's wavefronts, and small enough
        // to leave at least group_min_ groups per launch for load balance
        // (never below 8).
        int gla, glb, glc, gld;
        cls_ls(std::get<0>(kv.first), gla, glb);
        cls_ls(std::get<1>(kv.first), glc, gld);
        const int gfp32 = mixed && class_has_fp32(gla, glb, glc, gld) && lane_path(gla, glb, glc, gld, mixed) ? 1 : 0;
        const bool gdrow = class_has_fp32(gla, glb, glc, gld) &&
                           drow_plan(gla, glb, glc, gld, gfp32, (double)A.stat.prim_quartets, (double)A.stat.quartets);
        // a wavefront of the group kernel holds the classification flags of
        // its share of a group's tasks, one per lane: at most 64 x (its
        // workgroup's wavefronts) tasks per group
        const int gmax = 64 * (gdrow ? jk_group_waves(gla, glb, glc, gld, true) : 4);
        int cap = GROUP_MAX_TASKS > 0 ? std::min(GROUP_MAX_TASKS, gmax) : 0;
        if (cap <= 0) {
            const int rows = ncart(gla) + ncart(glb);
            // lists of cheap quartets (fewer than 4 primitive quartets each,
            // the uncontracted water-cluster lists) take groups up to 12
            // times as large: there the per-group row staging, zeroing and
            // flush weigh on quartets that cost little (measured on (H2O)32,
            // one stream: 46.4 -> 43.0 ms with 256-task groups; with the
            // density rows in LDS and up to 768-task groups 36.35 -> 36.03 ms)
            const bool cheap = (double)A.stat.prim_quartets < 4.0 * (double)A.stat.quartets;
            cap = std::min(gmax, std::max(8, (rows * nao_ + 63) / 64 * (cheap ? (gdrow ? 12 : 4) : 1)));
            // (per share when the build is dealt over shares: each share's
            // launch keeps group_min_ groups to balance over the device)
            cap = std::max(8, std::min(cap, (int)(A.tasks.size() / ((size_t)group_min_ * std::max(1, nshards_)))));
            cap = std::min(cap, gmax);
        }
        for (size_t t = 0; t < A.tasks.size();) {
            const int bra = A.tasks[t].bra;
            JKGroup G{bra, (int)t, 0, 0};
            double c = 0.0;
            while (t < A.tasks.size() && A.tasks[t].bra == bra && G.count < cap) {
                c += A.cost[t];
                G.count++;
                t++;
            }
            A.groups.push_back({c, G});
        }
        // heavy groups first for load balance
        std::stable_sort(A.groups.begin(), A.groups.end(),
                         [](const std::pair<double, JKGroup> &x, const std::pair<double, JKGroup> &y) {
                             return x.first > y.first;
                         });
        int la, lb, lc, ld;
        cls_ls(std::get<0>(kv.first), la, lb);
        cls_ls(std::get<1>(kv.first), lc, ld);
        if (nshards_ > 1) {
            // deal the groups (heaviest first, each to the least loaded
            // share; ties to the lowest index, so every share computes the
            // same deal) and keep this share's
            std::vector<double> load(nshards_, 0.0);
            std::vector<std::pair<double, JKGroup>> mine;
            for (auto &g : A.groups) {
                const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                load[s] += g.first;
                if (s == shard_) mine.push_back(g);
            }
            A.groups.swap(mine);
            // statistics of the kept tasks
            const double fp = flops_prim(la, lb, lc, ld);
            const int nf = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
            ClassLaunch S;
            for (auto &g : A.groups)
                for (int t = g.second.first; t < g.second.first + g.second.count; t++) {
                    const JKTask &T = A.tasks[t];
                    int b0, b1;
                    task_prims(T, pairs_[T.bra].npp, b0, b1);
                    const long long pq = (long long)T.nket * (b1 - b0) * pairs_[T.ket0].npp;
                    if (b0 == 0) S.quartets += T.nket;  // a split quartet counts once
                    S.prim_quartets += pq;
                    S.flops += (double)pq * fp + (b0 == 0 ? 12.0 * nf * T.nket : 0.0);
                    if (b0 == 0) S.lanes += 64;
                }
            stats_.quartets += S.quartets - A.stat.quartets;
            stats_.prim_quartets += S.prim_quartets - A.stat.prim_quartets;
            stats_.integrals += (long long)nf * (S.quartets - A.stat.quartets);
            stats_.flops += S.flops - A.stat.flops;
            A.stat.quartets = S.quartets;
            A.stat.prim_quartets = S.prim_quartets;
            A.stat.flops = S.flops;
            A.stat.lanes = S.lanes;
        }
        if (getenv("LIBRINT_DEBUG")) {
            double tot = 0.0, gmax = 0.0, tmax = 0.0;
            for (auto &g : A.groups) {
                tot += g.first;
                gmax = std::max(gmax, g.first);
                for (int t = g.second.first; t < g.second.first + g.second.count; t++) tmax = std::max(tmax, A.cost[t]);
            }
            fprintf(stderr, "librint: launch (%d%d|%d%d)%s tasks %zu groups %zu cost %.4g max group %.4g max task %.4g\n",
                    la, lb, lc, ld, std::get<2>(kv.first) ? " fp32" : "", A.tasks.size(), A.groups.size(), tot, gmax,
                    tmax);
        }
        ClassLaunch L = A.stat;
        L.la = la;
        L.lb = lb;
        L.lc = lc;
        L.ld = ld;
        L.fp32 = gfp32;
        L.drow_plan = gdrow ? 1 : 0;
        L.first = (int)tasks.size();
        L.gfirst = (int)groups.size();
        for (auto &g : A.groups) {
            JKGroup G = g.second;
            const int src = G.first;
            G.first = (int)tasks.size();
            for (int t = 0; t < G.count; t++) {
                tasks.push_back(A.tasks[src + t]);
                task_group.push_back((int)groups.size());
            }
            groups.push_back(G);
        }
        L.count = (int)tasks.size() - L.first;
        L.gcount = (int)groups.size() - L.gfirst;
        if (L.count) launches_.push_back(L);
    }
    stats_.launches = (int)launches_.size();
    dominant_ = 0;
    for (size_t k = 1; k < launches_.size(); k++)
        if (launches_[k].flops > launches_[dominant_].flops) dominant_ = (int)k;
    // launch order of a build: the dominant launch, then the others by
    // decreasing flops per task (fock_impl)
    launch_order_.assign(1, dominant_);
    for (int k = 0; k < (int)launches_.size(); k++)
        if (k != dominant_) launch_order_.push_back(k);
    std::stable_sort(launch_order_.begin() + 1, launch_order_.end(), [&](int x, int y) {
        return launches_[x].flops / std::max(1, launches_[x].count) > launches_[y].flops / std::max(1, launches_[y].count);
    });

    // Coulomb slabs: every launch of a register class stores its ket-side
    // Coulomb blocks per task (64 NFC NFD doubles) and a reduction sums the
    // slabs of the tasks that share a 64-ket window.  The slab memory is
    // bounded: with S doubles of slabs in all, a build keeps them all when
    // S <= B (one reduction once every launch has run: measured faster on
    // C16H34 than a reduction per launch on its stream), else each launch runs in
    // chunks of consecutive groups whose slabs fit B, each chunk reduced
    // right after it on its stream, with one B-sized buffer per stream.
    // B = LIBRINT_SLAB_MB MiB (default 2304: (H2O)32's largest launch, 2153
    // MiB of slabs, then runs as one chunk; measured with three streams
    // 35.47-35.50 ms against 35.76-35.85 with 1024), so the slabs take at most
    // max(min(S, B), B x streams) doubles.
    std::vector<JRedItem> items;
    std::vector<int> item_tasks;
    size_t budget = (size_t)2304 << 17;  // 2304 MiB of doubles
    if (const char *e = getenv("LIBRINT_SLAB_MB")) budget = std::max<size_t>(1, (size_t)atoll(e)) << 17;
    size_t total = 0;
    max_nfcd_ = 1;
    for (ClassLaunch &L : launches_) {
        L.jbase = -1;
        if (no_jslab_ || !class_has_fp32(L.la, L.lb, L.lc, L.ld) || !lane_path(L.la, L.lb, L.lc, L.ld, L.fp32))
            continue;
        L.jbase = (long long)total;
        total += (size_t)L.count * 64 * ncart(L.lc) * ncart(L.ld);
    }
    jslab_whole_ = total <= budget;
    chunks_.clear();
    for (ClassLaunch &L : launches_) {
        L.cfirst = (int)chunks_.size();
        const size_t per_task = 64 * (size_t)ncart(L.lc) * ncart(L.ld);
        if (L.jbase < 0) {
            chunks_.push_back(SlabChunk{L.gfirst, L.gcount, L.first, L.count, 0, 0, -1});
            L.ccount = 1;
            continue;
        }
        // consecutive groups (their tasks are consecutive) up to the budget
        for (int g = L.gfirst; g < L.gfirst + L.gcount;) {
            SlabChunk C{g, 0, groups[g].first, 0, (int)items.size(), 0, 0};
            size_t sz = 0;
            while (g < L.gfirst + L.gcount &&
                   (C.gcount == 0 || jslab_whole_ || sz + (size_t)groups[g].count * per_task <= budget)) {
                sz += (size_t)groups[g].count * per_task;
                C.tcount += groups[g].count;
                C.gcount++;
                g++;
            }
            C.jbase = jslab_whole_ ? L.jbase + (long long)(C.tfirst - L.first) * (long long)per_task : 0;
            std::map<int, std::vector<int>> win;
            for (int t = C.tfirst; t < C.tfirst + C.tcount; t++) win[tasks[t].ket0].push_back(t);
            // a window's task list is split into items of at most JRED_TASKS
            // tasks, so that the reduction has enough workgroups
            const int JRED_TASKS = 256;
            for (auto &kv : win) {
                const int cnt = (int)kv.second.size();
                item_tasks.insert(item_tasks.end(), kv.second.begin(), kv.second.end());
                for (int c0 = 0; c0 < cnt; c0 += JRED_TASKS) {
                    JRedItem it;
                    it.jbase = C.jbase;
                    it.tfirst = C.tfirst;
                    it.ket0 = kv.first;
                    it.nfc = ncart(L.lc);
                    it.nfd = ncart(L.ld);
                    it.lfirst = (int)item_tasks.size() - cnt + c0;
                    it.lcount = std::min(JRED_TASKS, cnt - c0);
                    items.push_back(it);
                    max_nfcd_ = std::max(max_nfcd_, it.nfc * it.nfd);
                }
            }
            C.icount = (int)items.size() - C.ifirst;
            chunks_.push_back(C);
        }
        L.ccount = (int)chunks_.size() - L.cfirst;
    }
    nitems_ = (int)items.size();
    dev_free(d_items_);
    dev_free(d_item_tasks_);
    d_items_ = dev_upload(items);
    d_item_tasks_ = dev_upload(item_tasks);
    // chunked mode: one buffer per stream, so that launches on different
    // streams never share slab memory
    slab_streams_ = jslab_whole_ ? 1 : std::max(1, nstreams_ > 0 ? nstreams_ : default_streams());
    jslot_ = jslab_whole_ ? total : budget;
    const size_t need = jslab_whole_ ? total : budget * slab_streams_;
    if (need > nJ_) {
        dev_free(d_J_);
        check(hipMalloc(&d_J_, sizeof(double) * need), "hipMalloc");
        nJ_ = need;
    }
    if (getenv("LIBRINT_DEBUG"))
        fprintf(stderr, "librint: coulomb slabs %zu of %zu doubles (%s), %zu chunks, %d reduction items\n", need, total,
                jslab_whole_ ? "one reduction" : "chunked", chunks_.size(), nitems_);
    for (hipEvent_t e : evc_) (void)hipEventDestroy(e);
    for (hipEvent_t e : evs_) (void)hipEventDestroy(e);
    evc_.assign(launches_.size(), nullptr);
    evs_.assign(launches_.size(), nullptr);
    for (auto &e : evc_) check(hipEventCreate(&e), "hipEventCreate");
    for (auto &e : evs_) check(hipEventCreate(&e), "hipEventCreate");
    for (hipEvent_t e : evr0_) (void)hipEventDestroy(e);
    for (hipEvent_t e : evr1_) (void)hipEventDestroy(e);
    evr0_.assign(chunks_.size(), nullptr);
    evr1_.assign(chunks_.size(), nullptr);
    red_rec_.assign(chunks_.size(), 0);
    for (auto &e : evr0_) check(hipEventCreate(&e), "hipEventCreate");
    for (auto &e : evr1_) check(hipEventCreate(&e), "hipEventCreate");
    dev_free(d_tasks_);
    d_tasks_ = dev_upload(tasks);
    dev_free(d_groups_);
    d_groups_ = dev_upload(groups);
    dev_free(d_task_group_);
    d_task_group_ = dev_upload(task_group);
    dev_free(d_flags_);
    dev_free(d_gmask_w_);
    dev_free(d_gmask_);
    check(hipMalloc(&d_flags_, std::max<size_t>(tasks.size(), 1)), "hipMalloc");
    check(hipMalloc(&d_gmask_w_, sizeof(unsigned int) * std::max<size_t>(groups.size(), 1)), "hipMalloc");
    check(hipMalloc(&d_gmask_, std::max<size_t>(groups.size(), 1)), "hipMalloc");
    ngroups_ = (int)groups.size();
    h_tasks_ = tasks;
    classified_ = false;
    dev_free(d_counters_);
    check(hipMalloc(&d_counters_, sizeof(int) * std::max<size_t>(chunks_.size(), 1)), "hipMalloc");
    check(hipMemset(d_counters_, 0, sizeof(int) * std::max<size_t>(chunks_.size(), 1)), "hipMemset");
}

void Context::set_precision(int mode, double tau_mix, double tau_skip)
{
    if (mode != 0 && mode != 1) throw std::invalid_argument("precision mode must be 0 (FP64) or 1 (mixed)");
    if (!(tau_mix >= 0) || !(tau_skip >= 0)) throw std::invalid_argument("thresholds must be >= 0");
    // the lists change only when the mixed kernels come or go; a new
    // threshold is applied by the next build
    const bool rebuild = (mode == 1 && tau_mix > 0.0) != (precision_ == 1 && tau_mix_ > 0.0);
    precision_ = mode;
    tau_mix_ = tau_mix;
    tau_skip_ = tau_skip;
    if (rebuild) build_tasks();
}

void Context::flag_counts(long long *q3, long long *p3)
{
    for (int f = 0; f < 3; f++) q3[f] = p3[f] = 0;
    std::vector<unsigned char> fl(h_tasks_.size(), 0);
    if ((classified_ || mixed_last_) && !fl.empty())
        check(hipMemcpy(fl.data(), d_flags_, fl.size(), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    for (const ClassLaunch &L : launches_) {
        for (int t = L.first; t < L.first + L.count; t++) {
            const JKTask &T = h_tasks_[t];
            // a mixed launch records 0 (FP64) or 1 (FP32) per evaluated task;
            // the other launches' tasks hold a classification flag (0 or 2)
            // or, without classification, nothing
            int f = 0;
            if (L.fp32 && L.variant == 1 && mixed_last_) f = fl[t];
            else if (classified_) f = fl[t] == 2 ? 2 : 0;
            int b0, b1;
            task_prims(T, pairs_[T.bra].npp, b0, b1);
            if (b0 == 0) q3[f] += T.nket;  // a split quartet counts once
            p3[f] += (long long)T.nket * (b1 - b0) * pairs_[T.ket0].npp;
        }
    }
}

void Context::fock_device(const double *D_dev, double *G_dev, hipStream_t st) { fock_impl(D_dev, G_dev, st, nullptr); }

size_t Context::scratch_doubles_needed() const
{
    const size_t nsh = shells_.size(), np = pairs_.size();
    const bool mixed = precision_ == 1 && tau_mix_ > 0.0;
    const bool dsh = tau_skip_ > 0.0 || mixed;
    // M, Dsh, the mixed-precision bounds, the counters
    size_t need = (size_t)nao_ * nao_ + (dsh ? nsh * nsh : 0) + (mixed ? nsh + 3 * np : 0) + launches_.size();
    bool hl = false;
    size_t hl_nf = 0;
    for (const ClassLaunch &L : launches_)
        if (hl_class(L.la, L.lb, L.lc, L.ld)) {
            hl = true;
            hl_nf = std::max<size_t>(hl_nf, (size_t)hl_shape(L.la, L.lb, L.lc, L.ld).nf);
        }
    return need + (hl ? hl_nf : 0);
}

void Context::fock_device_scratch(const double *D_dev, double *G_dev, hipStream_t st, double *scratch, size_t len)
{
    const size_t need = scratch_doubles_needed();
    if (need > len)
        throw std::invalid_argument("the Fock build needs " + std::to_string(need) +
                                    " doubles of scratch, more than the " + std::to_string(len) + " available");
    FockScratch sc;
    const size_t nn = (size_t)nao_ * nao_, nsh = shells_.size(), np = pairs_.size();
    const bool mixed = precision_ == 1 && tau_mix_ > 0.0;
    const bool dsh = tau_skip_ > 0.0 || mixed;
    double *p = scratch;
    sc.M = p;
    p += nn;
    sc.dsh = dsh ? p : nullptr;
    p += dsh ? nsh * nsh : 0;
    sc.mix = mixed ? p : nullptr;
    p += mixed ? nsh + 3 * np : 0;
    sc.counters = (int *)p;
    p += launches_.size();
    sc.hl = p;
    sc.hl_len = len - (size_t)(p - scratch);
    fock_impl(D_dev, G_dev, st, &sc);
}

void Context::fock_impl(const double *D_dev, double *G_dev, hipStream_t st, const FockScratch *sc)
{
    const size_t nn = (size_t)nao_ * nao_;
    double *M = nullptr;
    int *counters = d_counters_;
    if (sc) {
        // caller-provided scratch: nothing is allocated; M is cleared with a
        // kernel (no memset node if the caller captures the stream)
        M = sc->M;
        counters = sc->counters;
        check(launch_zero_doubles(M, nn, st), "zero M");
    } else {
        if (!d_M_) check(hipMalloc(&d_M_, sizeof(double) * nn), "hipMalloc");
        M = d_M_;
        check(hipMemsetAsync(d_M_, 0, sizeof(double) * nn, st), "hipMemsetAsync");
    }
    const DevBasis B = devb();
    check(launch_zero_counters(counters, sc ? (int)launches_.size() : (int)chunks_.size(), st), "zero counters");
    if (max_lds_ == 0) {
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&max_lds_, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "hipDeviceGetAttribute");
        if (getenv("LIBRINT_DEBUG")) fprintf(stderr, "librint: max LDS per block %d\n", max_lds_);
    }
    // Per-build task flags from the density's shell-block maxima: each task's
    // largest bound Q_ab Q_cd max(|D| over the quartet's six blocks) below
    // tau_skip: 2 (skipped); below tau_mix in mixed precision: 1 (the mixed
    // kernels run it with FP32 recurrences, the others in FP64); else 0
    const bool mixed = precision_ == 1 && tau_mix_ > 0.0;
    const bool classify = tau_skip_ > 0.0;
    const int nsh = (int)shells_.size();
    const int np = (int)pairs_.size();
    MixArgs mx{nullptr, nullptr, tau_mix_, d_flags_};
    if (classify || mixed) {
        double *dsh = sc ? sc->dsh : d_dsh_;
        if (!sc && !d_dsh_) {
            check(hipMalloc(&d_dsh_, sizeof(double) * (size_t)nsh * nsh), "hipMalloc");
            dsh = d_dsh_;
        }
        check(launch_dsh(D_dev, d_ao_off_, nsh, nao_, dsh, st), "dsh launch");
        if (classify)
            check(launch_classify(d_tasks_, (int)h_tasks_.size(), d_task_group_, ngroups_, B, D_dev, d_ao_off_, nsh,
                                  nao_, dsh, 0.0, tau_skip_, d_flags_, d_gmask_w_, d_gmask_, st),
                  "classify launch");
        if (mixed) {
            double *mb = sc ? sc->mix : d_mixb_;
            if (!sc && !d_mixb_) {
                check(hipMalloc(&d_mixb_, sizeof(double) * ((size_t)nsh + 3 * (size_t)np)), "hipMalloc");
                mb = d_mixb_;
            }
            double *drow = mb, *dpair = mb + nsh, *qd = dpair + np, *wqd = qd + np;
            check(launch_mix_bounds(dsh, nsh, B, np, drow, dpair, qd, wqd, st), "mixed-precision bounds launch");
            mx.Dpair = dpair;
            mx.Wqd = wqd;
        }
    }
    classified_ = classify;
    mixed_last_ = mixed;
    if (sc) std::fill(red_rec_.begin(), red_rec_.end(), 0);  // no slabs, no reductions
    const unsigned char *flags = classify ? d_flags_ : nullptr;
    const unsigned char *gmask = classify ? d_gmask_ : nullptr;
    check(hipEventRecord(ev0_, st), "hipEventRecord");
    // The launch with the most flops runs first and alone on st: it fills the
    // device by itself and its timing stays its own.  The others are then
    // dealt round-robin over st and ns - 1 extra streams, in order of
    // decreasing flops per task, so that the classes with long tasks and few
    // wavefronts per SIMD start early beside the others instead of running
    // alone at the end (measured: C16H34 1.53 -> 1.49 ms, (H2O)32 36.07 ->
    // 35.89 ms); they share nothing but M (atomics) and, in the per-launch
    // slab mode, their stream's slab buffer.
    const int nl = (int)launches_.size();
    const int nstreams = nstreams_ > 0 ? nstreams_ : default_streams();
    if (!sc) {
        std::vector<std::array<int, 4>> cl;
        for (const ClassLaunch &L : launches_) cl.push_back({L.la, L.lb, L.lc, L.ld});
        hl_scratch_for(cl);
    }
    // with caller scratch: one stream (nothing to create), no Coulomb slabs
    // (the ket Coulomb block goes to M by atomics)
    const bool slabs = sc == nullptr;
    int ns = std::min(nstreams, nl - 1);
    if (!jslab_whole_) ns = std::min(ns, slab_streams_);
    ns = std::max(ns, 1);
    if (sc) ns = 1;
    if (ns > 1) ensure_side_streams(ns - 1);
    stats_.streams = ns;
    for (int j = 0; j < nl; j++) {
        const int k = launch_order_[j];
        ClassLaunch &L = launches_[k];
        if (j == 1 && ns > 1) {
            check(hipEventRecord(evf_, st), "hipEventRecord");
            for (int i = 0; i < ns - 1; i++) check(hipStreamWaitEvent(side_[i], evf_, 0), "hipStreamWaitEvent");
        }
        // launches of the l > 2 classes share one scratch buffer: they all
        // stay on st
        const int slot = (j == 0 || hl_class(L.la, L.lb, L.lc, L.ld)) ? 0 : (j - 1) % ns;
        const hipStream_t sk = slot == 0 ? st : side_[slot - 1];
        double *jbuf = jslab_whole_ ? d_J_ : d_J_ + (size_t)slot * jslot_;
        check(hipEventRecord(evs_[k], sk), "hipEventRecord");
        const size_t lds = jk_group_lds(L.la, L.lb, L.lc, L.ld, L.fp32, nao_, false);
        const bool grouped = !no_group_ && lds <= (size_t)max_lds_;
        // the density-row staging planned with the task lists (drow_plan):
        // their group sizes assume its workgroup size
        const bool drow = grouped && L.drow_plan;
        L.drow = drow ? 1 : 0;
        const bool hl = hl_class(L.la, L.lb, L.lc, L.ld);
        L.variant = hl ? 3
                       : (grouped ? 1
                                  : (class_has_fp32(L.la, L.lb, L.lc, L.ld) && !class_split(L.la, L.lb, L.lc, L.ld) ? 0
                                                                                                                 : 2));
        if (hl) {
            const HLShape S = hl_shape(L.la, L.lb, L.lc, L.ld);
            check(launch_hl_jk(S, d_tasks_ + L.first, L.count, flags ? flags + L.first : nullptr, 3, B, D_dev, M, nao_,
                               sc ? sc->hl : hl_scratch(S), sc ? sc->hl_len : nhlscr_, sk),
                  "jk launch (l > 2)");
        } else {
            // the launch in its slab chunks (one when the slabs are kept
            // whole or absent); with caller scratch, no slabs: one launch
            // over every group with the ket Coulomb blocks added by atomics
            const int c_begin = sc ? 0 : L.cfirst, c_end = sc ? 1 : L.cfirst + L.ccount;
            for (int ci = c_begin; ci < c_end; ci++) {
                const SlabChunk C = sc ? SlabChunk{L.gfirst, L.gcount, L.first, L.count, 0, 0, -1} : chunks_[ci];
                int *ctr = sc ? counters + k : counters + ci;
                double *js = (slabs && C.jbase >= 0) ? jbuf + C.jbase : nullptr;
                if (grouped)
                    check(launch_jk_group(L.la, L.lb, L.lc, L.ld, L.fp32, d_groups_ + C.gfirst, C.gcount, d_tasks_,
                                          ctr, flags, gmask ? gmask + C.gfirst : nullptr, 3, B, D_dev, M, nao_, js,
                                          C.tfirst, drow, mx, sk),
                          "jk group launch");
                else
                    check(launch_jk(L.la, L.lb, L.lc, L.ld, L.fp32, d_tasks_ + C.tfirst, C.tcount, ctr,
                                    flags ? flags + C.tfirst : nullptr, 3, B, D_dev, M, nao_, js, 0, sk),
                          "jk launch");
                if (!sc) red_rec_[ci] = 0;
                if (js && !jslab_whole_ && C.icount) {
                    check(hipEventRecord(evr0_[ci], sk), "hipEventRecord");
                    check(launch_jreduce(d_items_ + C.ifirst, C.icount, ncart(L.lc) * ncart(L.ld), d_item_tasks_,
                                         d_tasks_, flags, 3, jbuf, B, M, nao_, sk),
                          "coulomb reduction launch");
                    check(hipEventRecord(evr1_[ci], sk), "hipEventRecord");
                    red_rec_[ci] = 1;
                }
            }
        }
        check(hipEventRecord(evc_[k], sk), "hipEventRecord");
    }
    for (int i = 0; i < ns - 1; i++) {
        check(hipEventRecord(evj_[i], side_[i]), "hipEventRecord");
        check(hipStreamWaitEvent(st, evj_[i], 0), "hipStreamWaitEvent");
    }
    if (slabs && jslab_whole_ && nitems_)
        check(launch_jreduce(d_items_, nitems_, max_nfcd_, d_item_tasks_, d_tasks_, flags, 3, d_J_, B, M, nao_, st),
              "coulomb reduction launch");
    check(hipEventRecord(ev1_, st), "hipEventRecord");
    check(launch_symmetrize(M, G_dev, nao_, st), "symmetrize launch");
}

// At least n extra streams (default priority, non-blocking) with one join
// event each, and the fork event.
void Context::ensure_side_streams(int n)
{
    while ((int)side_.size() < n) {
        hipStream_t q = nullptr;
        check(hipStreamCreateWithFlags(&q, hipStreamNonBlocking), "hipStreamCreate");
        side_.push_back(q);
        hipEvent_t e = nullptr;
        check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        evj_.push_back(e);
    }
    if (n > 0 && !evf_) check(hipEventCreateWithFlags(&evf_, hipEventDisableTiming), "hipEventCreate");
}

void Context::fock_host(const double *D, double *G)
{
    const size_t nn = (size_t)nao_ * nao_;
    if (!d_D_) check(hipMalloc(&d_D_, sizeof(double) * nn), "hipMalloc");
    if (!d_G_) check(hipMalloc(&d_G_, sizeof(double) * nn), "hipMalloc");
    check(hipMemcpy(d_D_, D, sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy H2D");
    fock_device(d_D_, d_G_, 0);
    check(hipMemcpy(G, d_G_, sizeof(double) * nn, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    check(hipEventElapsedTime(&stats_.last_kernel_ms, ev0_, ev1_), "hipEventElapsedTime");
    collect_class_times();
}

void Context::collect_class_times()
{
    for (size_t k = 0; k < launches_.size(); k++) {
        ClassLaunch &L = launches_[k];
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, evs_[k], evc_[k]) == hipSuccess) L.ms = ms;
        L.reduce_ms = 0.f;
        for (int ci = L.cfirst; ci < L.cfirst + L.ccount && ci < (int)red_rec_.size(); ci++) {
            float r = 0.f;
            if (red_rec_[ci] && hipEventElapsedTime(&r, evr0_[ci], evr1_[ci]) == hipSuccess) L.reduce_ms += r;
        }
    }
}

void Context::tensor_host(double *R)
{
    const size_t n = (size_t)nao_;
    const size_t n4 = n * n * n * n;
    double *dR = nullptr;
    check(hipMalloc(&dR, sizeof(double) * n4), "hipMalloc");
    try {
        tensor_device(dR);
    } catch (...) {
        dev_free(dR);
        throw;
    }
    check(hipMemcpy(R, dR, sizeof(double) * n4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    dev_free(dR);
}

void Context::tensor_device(double *dR)
{
    const size_t n = (size_t)nao_;
    const size_t n4 = n * n * n * n;
    check(hipMemset(dR, 0, sizeof(double) * n4), "hipMemset");
    // every canonical pair of pairs, no screening (exact tensor)
    const int np = (int)pairs_.size();
    std::map<std::pair<int, int>, std::vector<ListTask>> by_cls;
    for (int p = 0; p < np; p++)
        for (int q = 0; q <= p; q++) {
            int a = p, b = q;
            if (pairs_[a].cls < pairs_[b].cls) std::swap(a, b);
            by_cls[{pairs_[a].cls, pairs_[b].cls}].push_back(ListTask{a, b, 0});
        }
    const DevBasis B = devb();
    std::vector<ListTask *> bufs;
    hl_scratch_for(list_classes(by_cls));
    for (auto &kv : by_cls) {
        int la, lb, lc, ld;
        cls_ls(kv.first.first, la, lb);
        cls_ls(kv.first.second, lc, ld);
        ListTask *dt = dev_upload(kv.second);
        bufs.push_back(dt);
        check(list_launch(la, lb, lc, ld, dt, (int)kv.second.size(), B, dR, 2, 0), "tensor launch");
    }
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    for (auto p : bufs) dev_free(p);
}

size_t Context::shells_host(const int *quartets, int nq, double *out)
{
    // A generally contracted input shell is a run of nctr virtual segmented
    // shells.  Each requested (I,J,K,L) expands into one sub-block per
    // contraction combination (c0,c1,c2,c3); each sub-block maps to a
    // canonical stored pair of pairs, is computed in stored orientation and
    // permuted into the requested block at I = f + nf*c (cint2e_cart layout,
    // reference src/cint2e.rs:3128-3152).
    struct Req {
        int bra, ket;
        bool swap_bk;
        long long off_dev;
        long long off_out;
        int vs[4];   // virtual shells of the sub-block
        int nf[4];   // cartesian functions per virtual shell
        int ctr[4];  // contraction index of each slot
        int dims[4]; // full block dims (nf * nctr)
    };
    std::vector<int> cum(bas_.size() / 8 + 1, 0);  // first virtual shell of each input shell
    {
        int v = 0;
        for (size_t s = 0; s < bas_.size() / 8; s++) {
            cum[s] = v;
            v += bas_[8 * s + 3];
        }
        cum[bas_.size() / 8] = v;
    }
    auto pidx = [&](int x, int y) {
        if (x < y) std::swap(x, y);
        return pair_of_ij_[(size_t)x * (x + 1) / 2 + y];
    };
    std::vector<Req> reqs;
    std::map<std::pair<int, int>, std::vector<ListTask>> by_cls;
    long long dev_total = 0, out_total = 0;
    for (int t = 0; t < nq; t++) {
        const int *s = quartets + 4 * t;
        for (int a = 0; a < 4; a++)
            if (s[a] < 0 || (size_t)s[a] >= bas_.size() / 8) throw std::invalid_argument("shell index out of range");
        int nctr[4], dims[4];
        for (int a = 0; a < 4; a++) {
            nctr[a] = bas_[8 * s[a] + 3];
            dims[a] = shells_[cum[s[a]]].nf * nctr[a];
        }
        for (int c3 = 0; c3 < nctr[3]; c3++)
            for (int c2 = 0; c2 < nctr[2]; c2++)
                for (int c1 = 0; c1 < nctr[1]; c1++)
                    for (int c0 = 0; c0 < nctr[0]; c0++) {
                        Req r;
                        const int cc[4] = {c0, c1, c2, c3};
                        for (int a = 0; a < 4; a++) {
                            r.ctr[a] = cc[a];
                            r.vs[a] = cum[s[a]] + cc[a];
                            r.nf[a] = shells_[r.vs[a]].nf;
                            r.dims[a] = dims[a];
                        }
                        r.bra = pidx(r.vs[0], r.vs[1]);
                        r.ket = pidx(r.vs[2], r.vs[3]);
                        r.off_out = out_total;
                        r.swap_bk = false;
                        r.off_dev = -1;
                        if (r.bra >= 0 && r.ket >= 0) {
                            if (pairs_[r.bra].cls < pairs_[r.ket].cls) {
                                std::swap(r.bra, r.ket);
                                r.swap_bk = true;
                            }
                            r.off_dev = dev_total;
                            dev_total += (long long)r.nf[0] * r.nf[1] * r.nf[2] * r.nf[3];
                            by_cls[{pairs_[r.bra].cls, pairs_[r.ket].cls}].push_back(
                                ListTask{r.bra, r.ket, r.off_dev});
                        }
                        reqs.push_back(r);
                    }
        out_total += (long long)dims[0] * dims[1] * dims[2] * dims[3];
    }
    std::vector<double> blocks(std::max<long long>(dev_total, 1));
    if (dev_total) {
        double *dout = nullptr;
        check(hipMalloc(&dout, sizeof(double) * dev_total), "hipMalloc");
        const DevBasis B = devb();
        std::vector<ListTask *> bufs;
        hl_scratch_for(list_classes(by_cls));
        for (auto &kv : by_cls) {
            int la, lb, lc, ld;
            cls_ls(kv.first.first, la, lb);
            cls_ls(kv.first.second, lc, ld);
            ListTask *dt = dev_upload(kv.second);
            bufs.push_back(dt);
            check(list_launch(la, lb, lc, ld, dt, (int)kv.second.size(), B, dout, 0, 0), "shells launch");
        }
        check(hipMemcpy(blocks.data(), dout, sizeof(double) * dev_total, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        for (auto p : bufs) dev_free(p);
        dev_free(dout);
    }
    for (const Req &r : reqs) {
        const int d0 = r.dims[0], d1 = r.dims[1], d2 = r.dims[2];
        double *o = out + r.off_out;
        const int n0 = r.nf[0], n1 = r.nf[1], n2 = r.nf[2], n3 = r.nf[3];
        const int o0 = n0 * r.ctr[0], o1 = n1 * r.ctr[1], o2 = n2 * r.ctr[2], o3 = n3 * r.ctr[3];
        if (r.off_dev < 0) {  // no surviving primitive pair: the sub-block is zero
            for (int l = 0; l < n3; l++)
                for (int k = 0; k < n2; k++)
                    for (int j = 0; j < n1; j++)
                        for (int i = 0; i < n0; i++) o[(o0 + i) + d0 * ((o1 + j) + d1 * ((o2 + k) + d2 * (o3 + l)))] = 0.0;
            continue;
        }
        // stored block: pair X = (sa, sb), pair Y = (sc, sd), a fastest
        const PairH &X = pairs_[r.bra], &Y = pairs_[r.ket];
        const int na = shells_[X.sa].nf, nbf = shells_[X.sb].nf, nc = shells_[Y.sa].nf;
        const int b0 = r.swap_bk ? 2 : 0, k0 = r.swap_bk ? 0 : 2;
        for (int l = 0; l < n3; l++)
            for (int k = 0; k < n2; k++)
                for (int j = 0; j < n1; j++)
                    for (int i = 0; i < n0; i++) {
                        // component of each stored slot: the stored bra pair
                        // (sa, sb) comes from requested slots b0, b0 + 1
                        // (in either order), the ket pair from k0, k0 + 1
                        int fi[4];
                        const int fv[4] = {i, j, k, l};
                        if (r.vs[b0] == X.sa && r.vs[b0 + 1] == X.sb) {
                            fi[0] = fv[b0];
                            fi[1] = fv[b0 + 1];
                        } else {
                            fi[0] = fv[b0 + 1];
                            fi[1] = fv[b0];
                        }
                        if (r.vs[k0] == Y.sa && r.vs[k0 + 1] == Y.sb) {
                            fi[2] = fv[k0];
                            fi[3] = fv[k0 + 1];
                        } else {
                            fi[2] = fv[k0 + 1];
                            fi[3] = fv[k0];
                        }
                        o[(o0 + i) + d0 * ((o1 + j) + d1 * ((o2 + k) + d2 * (o3 + l)))] =
                            blocks[r.off_dev + fi[0] + na * (fi[1] + nbf * (fi[2] + nc * fi[3]))];
                    }
    }
    return (size_t)out_total;
}

void Context::build_grad_tasks()
{
    // Ket-major list of the quartets of build_tasks (bins X >= Y, Q_i Q_j >
    // tau, j <= i inside one bin): for each ket pair j of bin Y the bras of
    // bin X form a prefix in Q order, cut into windows of 64 (grad_kets_ for
    // the classes whose Fock tasks take short windows).
    glaunch1_.clear();
    if (d_gtasks1_) check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    dev_free(d_gtasks1_);
    const int nb = (int)bins_.size();
    std::map<std::pair<int, int>, std::vector<JKTask>> acc;  // (bra class, ket class)
    for (int X = 0; X < nb; X++) {
        const Bin &bx = bins_[X];
        for (int Y = 0; Y <= X; Y++) {
            const Bin &by = bins_[Y];
            std::vector<JKTask> &v = acc[std::make_pair(bx.cls, by.cls)];
            for (int j = 0; j < by.count; j++) {
                const int k = by.first + j;
                const double Qk = pairs_[k].Q;
                int lo = 0, hi = bx.count;
                while (lo < hi) {
                    const int mid = (lo + hi) / 2;
                    if (pairs_[bx.first + mid].Q * Qk > tau_) lo = mid + 1;
                    else hi = mid;
                }
                int la, lb, lc, ld;
                cls_ls(bx.cls, la, lb);
                cls_ls(by.cls, lc, ld);
                const int win = class_has_fp32(la, lb, lc, ld) && lane_path(la, lb, lc, ld, 0) ? 64 : grad_kets_;
                for (int i0 = X == Y ? j : 0; i0 < lo; i0 += win)
                    v.push_back(JKTask{k, bx.first + i0, std::min(win, lo - i0), 0});
            }
        }
    }
    std::vector<JKTask> all;
    for (auto &kv : acc) {
        if (kv.second.empty()) continue;
        GradLaunch G;
        cls_ls(kv.first.first, G.la, G.lb);
        cls_ls(kv.first.second, G.lc, G.ld);
        G.first = (int)all.size();
        G.count = (int)kv.second.size();
        all.insert(all.end(), kv.second.begin(), kv.second.end());
        glaunch1_.push_back(G);
    }
    if (!all.empty()) d_gtasks1_ = dev_upload(all);
    // bra side: the Fock tasks, with those of the chunked classes (64-ket
    // windows in the Fock build) cut into the short windows the adjoint's
    // one-ket-at-a-time tasks balance with
    glaunch0_.clear();
    dev_free(d_gtasks0_);
    std::vector<JKTask> all0;
    std::vector<int> bin_of(pairs_.size(), -1);
    for (int b = 0; b < nb; b++)
        for (int k = bins_[b].first; k < bins_[b].first + bins_[b].count; k++) bin_of[k] = b;
    for (const ClassLaunch &L : launches_) {
        GradLaunch G{L.la, L.lb, L.lc, L.ld, (int)all0.size(), 0};
        const bool lane = class_has_fp32(L.la, L.lb, L.lc, L.ld) && lane_path(L.la, L.lb, L.lc, L.ld, L.fp32);
        if (lane && class_split(L.la, L.lb, L.lc, L.ld)) {
            for (int t = L.first; t < L.first + L.count; t++) {
                const JKTask &T = h_tasks_[t];
                for (int k0 = 0; k0 < T.nket; k0 += grad_kets_)
                    all0.push_back(JKTask{T.bra, T.ket0 + k0, std::min(grad_kets_, T.nket - k0), T.prims});
            }
        } else if (!lane && grad_kets_ != team_kets_) {
            // the classes whose Fock tasks take team_kets_ windows: windows
            // of grad_kets_, cut from those tasks and joined where they
            // continue one another over the kets of one bin (a task never
            // spans two bins: the kernel's primitive loops are uniform)
            for (int t = L.first; t < L.first + L.count; t++) {
                const JKTask &T = h_tasks_[t];
                for (int k0 = 0; k0 < T.nket; k0 += grad_kets_) {
                    const JKTask w{T.bra, T.ket0 + k0, std::min(grad_kets_, T.nket - k0), T.prims};
                    JKTask *last = (int)all0.size() > G.first ? &all0.back() : nullptr;
                    if (last && last->bra == w.bra && last->prims == w.prims && last->ket0 + last->nket == w.ket0 &&
                        bin_of[last->ket0] == bin_of[w.ket0] && last->nket + w.nket <= grad_kets_)
                        last->nket += w.nket;
                    else
                        all0.push_back(w);
                }
            }
        }
        G.count = (int)all0.size() - G.first;
        glaunch0_.push_back(G);
    }
    if (!all0.empty()) d_gtasks0_ = dev_upload(all0);
    gtasks_valid_ = true;
}

void Context::grad2e(const double *P, double *gatom, double *denv, size_t env_l)
{
    if (nshards_ > 1) throw std::invalid_argument("analytic gradients need the whole build (nshards = 1)");
    // per-shell primitive slots: exponent and coefficient derivatives are
    // accumulated per (shell, primitive) and folded onto env entries here
    const int ns = (int)shells_.size();
    if (!d_pp_idx_) {
        std::vector<int> slot(ns);
        std::vector<int2> info(ns);
        std::vector<double> cf;
        int nslot = 0;
        for (int s = 0; s < ns; s++) {
            slot[s] = nslot;
            info[s] = make_int2(shells_[s].atom, nslot);
            for (int ip = 0; ip < shells_[s].nprim; ip++) cf.push_back(env_[shells_[s].pcoef + ip]);
            nslot += shells_[s].nprim;
        }
        nslot_ = nslot;
        std::vector<int2> idx(pp_.size());
        for (size_t k = 0; k < pp_.size(); k++) idx[k] = make_int2(pp_idx_[2 * k], pp_idx_[2 * k + 1]);
        d_pp_idx_ = dev_upload(idx);
        d_shell_slot_ = dev_upload(info);
        d_slot_coef_ = dev_upload(cf);
    }
    if (!gtasks_valid_) build_grad_tasks();
    const size_t nn = (size_t)nao_ * nao_;
    const int na = natm();
    if (!d_D_) check(hipMalloc(&d_D_, sizeof(double) * nn), "hipMalloc");
    check(hipMemcpy(d_D_, P, sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy H2D");
    double *dg = nullptr;
    const size_t nout = 3 * (size_t)na + 2 * (size_t)nslot_;
    check(hipMalloc(&dg, sizeof(double) * nout), "hipMalloc");
    check(hipMemset(dg, 0, sizeof(double) * nout), "hipMemset");
    const DevBasis B = devb();
    // The class launches of both sides are dealt round-robin over the null
    // stream and nst - 1 extra streams (LIBRINT_GRAD_STREAMS, default 4): they
    // share only the atomically updated outputs, and the small launches and
    // launch tails overlap.
    int nst = 4;
    if (const char *e = getenv("LIBRINT_GRAD_STREAMS")) nst = std::max(1, std::min(8, atoi(e)));
    ensure_side_streams(nst - 1);
    check(hipEventRecord(ev0_, 0), "hipEventRecord");
    if (nst > 1) {
        check(hipEventRecord(evf_, 0), "hipEventRecord");
        for (int i = 0; i < nst - 1; i++) check(hipStreamWaitEvent(side_[i], evf_, 0), "hipStreamWaitEvent");
    }
    int k = 0;
    auto stream_of = [&](int j) { return (j % nst) == 0 ? (hipStream_t)0 : side_[(j % nst) - 1]; };
    // the f/g classes (run-time shaped kernel) share the scratch of one
    // weight block per workgroup: they all run on the null stream, sized
    // for the largest class before any launch
    {
        std::vector<std::array<int, 4>> cl;
        for (const ClassLaunch &L : launches_) cl.push_back({L.la, L.lb, L.lc, L.ld});
        hl_scratch_for(cl);
    }
    auto grad_launch = [&](int la, int lb, int lc, int ld, int side, const JKTask *tl, int cnt) {
        if (hl_class(la, lb, lc, ld)) {
            const HLShape S = hl_shape(la, lb, lc, ld);
            check(launch_hl_grad(la, lb, lc, ld, side, tl, cnt, B, d_pp_idx_, d_shell_slot_, d_slot_coef_, d_D_,
                                 nao_, dg, dg + 3 * na, hl_scratch(S), nhlscr_, (hipStream_t)0),
                  "grad launch (l > 2)");
        } else {
            check(launch_grad(la, lb, lc, ld, side, tl, cnt, B, d_pp_idx_, d_shell_slot_, d_slot_coef_, d_D_, nao_,
                              dg, dg + 3 * na, stream_of(k++)),
                  "grad launch");
        }
    };
    for (size_t k = 0; k < launches_.size(); k++) {
        const ClassLaunch &L = launches_[k];
        const GradLaunch &G = glaunch0_[k];
        if (G.count)
            grad_launch(L.la, L.lb, L.lc, L.ld, 0, d_gtasks0_ + G.first, G.count);
        else
            grad_launch(L.la, L.lb, L.lc, L.ld, 0, d_tasks_ + L.first, L.count);
    }
    for (const GradLaunch &G : glaunch1_) grad_launch(G.la, G.lb, G.lc, G.ld, 1, d_gtasks1_ + G.first, G.count);
    for (int i = 0; i < nst - 1; i++) {
        check(hipEventRecord(evj_[i], side_[i]), "hipEventRecord");
        check(hipStreamWaitEvent(0, evj_[i], 0), "hipStreamWaitEvent");
    }
    check(hipEventRecord(ev1_, 0), "hipEventRecord");
    std::vector<double> h(nout);
    check(hipMemcpy(h.data(), dg, sizeof(double) * nout, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    check(hipEventElapsedTime(&stats_.last_grad_ms, ev0_, ev1_), "hipEventElapsedTime");
    dev_free(dg);
    for (int k = 0; k < 3 * na; k++) gatom[k] = h[k];
    std::fill(denv, denv + env_l, 0.0);
    const double *dp = h.data() + 3 * na;
    int slot = 0;
    for (int s = 0; s < ns; s++) {
        const Shell &S = shells_[s];
        for (int ip = 0; ip < S.nprim; ip++, slot++) {
            if ((size_t)(S.pexp + ip) < env_l) denv[S.pexp + ip] += dp[2 * slot];
            if ((size_t)(S.pcoef + ip) < env_l) denv[S.pcoef + ip] += dp[2 * slot + 1];
        }
    }
}

std::vector<double> Context::schwarz_pairs() const
{
    const int ns = (int)shells_.size();
    std::vector<double> q((size_t)ns * (ns + 1) / 2, 0.0);
    for (size_t k = 0; k < q.size(); k++)
        if (pair_of_ij_[k] >= 0) q[k] = pairs_[pair_of_ij_[k]].Q;
    return q;
}

}  // namespace librint
