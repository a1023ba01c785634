// extern "C" entry points (include/librint.h).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "../../include/librint.h"
#include "context.h"
#include "int1e.h"
#include "scf.h"

using namespace librint;

struct librint_ctx {
    std::unique_ptr<Context> ctx;
    Int1eInput one;
};

static thread_local std::string g_err;

static void set_err(const std::string &m)
{
    g_err = m;
    std::fprintf(stderr, "librint: %s\n", m.c_str());
}

static double *alloc_out(size_t n)
{
    double *p = (double *)std::malloc(sizeof(double) * (n ? n : 1));
    if (!p) throw std::bad_alloc();
    return p;
}

static bool coord_ok(int coord)
{
    if (coord == 0) return true;
    if (coord == 1) set_err("spherical (coord=1) output is not provided by the GPU path yet");
    else set_err("invalid coord code " + std::to_string(coord));
    return false;
}

extern "C" {

void free_c(double *ptr, size_t len)
{
    (void)len;
    std::free(ptr);
}

const char *librint_last_error(void) { return g_err.c_str(); }

double *int1e_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, int coord, int typec)
{
    (void)env_l;
    try {
        if (!coord_ok(coord)) return nullptr;
        if (typec < 0 || typec > 2) {
            set_err("invalid typec code " + std::to_string(typec));
            return nullptr;
        }
        Int1eInput in = make_int1e_input(atm, atm_l, bas, bas_l, env);
        double *R = alloc_out((size_t)in.nao * in.nao);
        int1e_matrix(in, typec, R);
        return R;
    } catch (std::exception &e) {
        set_err(e.what());
        return nullptr;
    }
}

double *int2e_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, int coord)
{
    try {
        if (!coord_ok(coord)) return nullptr;
        Context ctx(atm, atm_l, bas, bas_l, env, env_l);
        const size_t n = (size_t)ctx.nao();
        double *R = alloc_out(n * n * n * n);
        ctx.tensor_host(R);
        return R;
    } catch (std::exception &e) {
        set_err(e.what());
        return nullptr;
    }
}

static ScfOut run_ref_scf(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, size_t nelec,
                          int imax, double conv)
{
    Context ctx(atm, atm_l, bas, bas_l, env, env_l);
    ctx.set_schwarz(0.0);
    Int1eInput in = make_int1e_input(atm, atm_l, bas, bas_l, env);
    ScfOptions o;
    return scf_run(ctx, in, (int)nelec, imax, conv, o);
}

double *density_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, size_t nelec, int imax,
                  double conv)
{
    try {
        ScfOut r = run_ref_scf(atm, atm_l, bas, bas_l, env, env_l, nelec, imax, conv);
        if (r.status != 0) {
            set_err(r.msg);
            return nullptr;
        }
        double *P = alloc_out(r.P.size());
        std::memcpy(P, r.P.data(), sizeof(double) * r.P.size());
        return P;
    } catch (std::exception &e) {
        set_err(e.what());
        return nullptr;
    }
}

double energy_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l)
{
    try {
        Context ctx(atm, atm_l, bas, bas_l, env, env_l);
        ctx.set_schwarz(0.0);
        const size_t n = (size_t)ctx.nao();
        if (P_l != n * n) {
            set_err("P has the wrong length");
            return NAN;
        }
        Int1eInput in = make_int1e_input(atm, atm_l, bas, bas_l, env);
        std::vector<double> H(n * n), V(n * n);
        int1e_matrix(in, 1, H.data());
        int1e_matrix(in, 2, V.data());
        for (size_t i = 0; i < n * n; i++) H[i] += V[i];
        return energy_fast(ctx, in, P, H.data());
    } catch (std::exception &e) {
        set_err(e.what());
        return NAN;
    }
}

double scf_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, size_t nelec, int imax,
             double conv)
{
    try {
        ScfOut r = run_ref_scf(atm, atm_l, bas, bas_l, env, env_l, nelec, imax, conv);
        if (r.status != 0) {
            set_err(r.msg);
            return NAN;
        }
        return r.E;
    } catch (std::exception &e) {
        set_err(e.what());
        return NAN;
    }
}

#define LR_NOT_PROVIDED(name)                                                                                          \
    set_err(#name ": basis-parameter gradients are not provided by this library yet");                                 \
    return nullptr;

double *dS_u(int *, size_t, int *, size_t, double *, size_t) { LR_NOT_PROVIDED(dS_u) }
double *grad_c(int *, size_t, int *, size_t, double *, size_t, double *, size_t) { LR_NOT_PROVIDED(grad_c) }
double *dS_c(int *, size_t, int *, size_t, double *, size_t, double *, size_t) { LR_NOT_PROVIDED(dS_c) }
double *dHcore_c(int *, size_t, int *, size_t, double *, size_t, double *, size_t) { LR_NOT_PROVIDED(dHcore_c) }
double *dR_c(int *, size_t, int *, size_t, double *, size_t, double *, size_t) { LR_NOT_PROVIDED(dR_c) }
double *danalytical_c(int *, size_t, int *, size_t, double *, size_t, double *, size_t) { LR_NOT_PROVIDED(danalytical_c) }
double *denergy_c(int *, size_t, int *, size_t, double *, size_t, double *, size_t) { LR_NOT_PROVIDED(denergy_c) }

double *fock_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l,
               int coord, int precision, double schwarz_tol)
{
    try {
        if (!coord_ok(coord)) return nullptr;
        if (precision != 0) {
            set_err("precision mode " + std::to_string(precision) + " is not available");
            return nullptr;
        }
        Context ctx(atm, atm_l, bas, bas_l, env, env_l);
        ctx.set_schwarz(schwarz_tol);
        const size_t n = (size_t)ctx.nao();
        if (P_l != n * n) {
            set_err("P has the wrong length");
            return nullptr;
        }
        double *G = alloc_out(n * n);
        ctx.fock_host(P, G);
        return G;
    } catch (std::exception &e) {
        set_err(e.what());
        return nullptr;
    }
}

librint_ctx *librint_ctx_create(const int *atm, size_t atm_l, const int *bas, size_t bas_l, const double *env,
                                size_t env_l)
{
    try {
        auto *c = new librint_ctx;
        c->ctx.reset(new Context(atm, atm_l, bas, bas_l, env, env_l));
        c->one = make_int1e_input(atm, atm_l, bas, bas_l, env);
        return c;
    } catch (std::exception &e) {
        set_err(e.what());
        return nullptr;
    }
}

void librint_ctx_destroy(librint_ctx *ctx) { delete ctx; }

int librint_ctx_nao(const librint_ctx *ctx) { return ctx ? ctx->ctx->nao() : -1; }

int librint_ctx_set_schwarz(librint_ctx *ctx, double tol)
{
    try {
        if (!ctx || !(tol >= 0)) {
            set_err("invalid context or Schwarz threshold");
            return -1;
        }
        ctx->ctx->set_schwarz(tol);
        return 0;
    } catch (std::exception &e) {
        set_err(e.what());
        return -1;
    }
}

int librint_ctx_set_precision(librint_ctx *ctx, int mode, double tau_mix, double tau_skip, double dmax)
{
    try {
        if (!ctx) {
            set_err("null context");
            return -1;
        }
        ctx->ctx->set_precision(mode, tau_mix, tau_skip, dmax);
        return 0;
    } catch (std::exception &e) {
        set_err(e.what());
        return -1;
    }
}

int librint_ctx_flag_counts(librint_ctx *ctx, double *out6)
{
    try {
        if (!ctx || !out6) {
            set_err("null argument");
            return -1;
        }
        long long q[3], p[3];
        ctx->ctx->flag_counts(q, p);
        for (int i = 0; i < 3; i++) {
            out6[i] = (double)q[i];
            out6[3 + i] = (double)p[i];
        }
        return 0;
    } catch (std::exception &e) {
        set_err(e.what());
        return -1;
    }
}

int librint_fock(librint_ctx *ctx, const double *P, double *G)
{
    try {
        if (!ctx || !P || !G) {
            set_err("null argument");
            return -1;
        }
        ctx->ctx->fock_host(P, G);
        return 0;
    } catch (std::exception &e) {
        set_err(e.what());
        return -1;
    }
}

int librint_fock_device(librint_ctx *ctx, const double *P_dev, double *G_dev, void *stream)
{
    try {
        if (!ctx || !P_dev || !G_dev) {
            set_err("null argument");
            return -1;
        }
        ctx->ctx->fock_device(P_dev, G_dev, (hipStream_t)stream);
        return 0;
    } catch (std::exception &e) {
        set_err(e.what());
        return -1;
    }
}

size_t librint_int2e_shells(librint_ctx *ctx, const int *quartets, int nq, double *out)
{
    try {
        if (!ctx || (nq > 0 && !quartets)) {
            set_err("null argument");
            return (size_t)-1;
        }
        const std::vector<int> &bas = ctx->ctx->bas();
        size_t total = 0;
        for (int t = 0; t < nq; t++) {
            size_t d = 1;
            for (int a = 0; a < 4; a++) {
                const int s = quartets[4 * t + a];
                if (s < 0 || (size_t)s >= bas.size() / 8) {
                    set_err("shell index out of range");
                    return (size_t)-1;
                }
                const int l = bas[8 * s + 1];
                d *= (size_t)(l + 1) * (l + 2) / 2 * bas[8 * s + 3];
            }
            total += d;
        }
        if (!out) return total;
        return ctx->ctx->shells_host(quartets, nq, out);
    } catch (std::exception &e) {
        set_err(e.what());
        return (size_t)-1;
    }
}

int librint_schwarz(librint_ctx *ctx, double *out)
{
    if (!ctx || !out) {
        set_err("null argument");
        return -1;
    }
    std::vector<double> q = ctx->ctx->schwarz_pairs();
    std::memcpy(out, q.data(), sizeof(double) * q.size());
    return 0;
}

int librint_ctx_stats(const librint_ctx *ctx, double *out, int n)
{
    if (!ctx || !out) return -1;
    const Stats &s = ctx->ctx->stats();
    const double v[8] = {(double)s.quartets, (double)s.prim_quartets, (double)s.integrals, s.flops,
                         (double)s.candidates, (double)s.launches, (double)s.last_kernel_ms, ctx->ctx->schwarz()};
    int k = 0;
    for (; k < n && k < 8; k++) out[k] = v[k];
    return k;
}

int librint_ctx_class_stats(const librint_ctx *ctx, double *out, int n)
{
    if (!ctx) return -1;
    const auto &L = ctx->ctx->launches();
    if (!out) return (int)L.size();
    int k = 0;
    for (; k < (int)L.size() && k < n; k++) {
        const ClassLaunch &c = L[k];
        const double v[10] = {(double)c.la, (double)c.lb, (double)c.lc, (double)c.ld, (double)c.quartets,
                              (double)c.prim_quartets, c.flops, (double)c.lanes, (double)c.ms, (double)c.count};
        for (int j = 0; j < 10; j++) out[10 * k + j] = v[j];
    }
    return (int)L.size();
}

double librint_scf(librint_ctx *ctx, int nelec, int imax, double conv, int guess, int diis, double *P_inout,
                   int *iters)
{
    try {
        if (!ctx) {
            set_err("null context");
            return NAN;
        }
        ScfOptions o;
        o.guess = guess;
        o.diis = diis;
        o.P0 = (guess == 2) ? P_inout : nullptr;
        ScfOut r = scf_run(*ctx->ctx, ctx->one, nelec, imax, conv, o);
        if (iters) *iters = r.iters;
        if (P_inout && !r.P.empty()) std::memcpy(P_inout, r.P.data(), sizeof(double) * r.P.size());
        if (r.status != 0) {
            set_err(r.msg);
            return NAN;
        }
        return r.E;
    } catch (std::exception &e) {
        set_err(e.what());
        return NAN;
    }
}

}  // extern "C"
