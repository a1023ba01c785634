// extern "C" entry points (include/librint.h).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/librint.h"
#include "context.h"
#include "int1e.h"
#include "cart2sph.h"
#include "int1e_gpu.h"
#include "mlp.h"
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
    if (coord == 0 || coord == 1) return true;
    set_err("invalid coord code " + std::to_string(coord));
    return false;
}

// device buffer released on scope exit
struct DevMem {
    double *p = nullptr;
    explicit DevMem(size_t n) { check(hipMalloc(&p, sizeof(double) * std::max<size_t>(n, 1)), "hipMalloc"); }
    ~DevMem()
    {
        if (p) (void)hipFree(p);
    }
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
};

// spherical form of a cartesian ndim-index array on the device, copied to host
static double *to_spherical(int ndim, const int *bas, size_t bas_l, const double *dcart, int ncart)
{
    int nc = 0;
    std::vector<C2SRow> rows = c2s_rows(bas, bas_l, &nc);
    if (nc != ncart) throw std::runtime_error("cartesian count mismatch in the spherical transform");
    const size_t ns = rows.size();
    const size_t m = ndim == 2 ? ns * ns : ns * ns * ns * ns;
    C2SRow *drows = nullptr;
    check(hipMalloc(&drows, sizeof(C2SRow) * std::max<size_t>(ns, 1)), "hipMalloc");
    struct RowsFree {
        C2SRow *p;
        ~RowsFree() { (void)hipFree(p); }
    } rf{drows};
    check(hipMemcpy(drows, rows.data(), sizeof(C2SRow) * ns, hipMemcpyHostToDevice), "hipMemcpy H2D");
    DevMem dsph(m);
    check(launch_c2s(ndim, drows, (int)ns, dcart, ncart, dsph.p, 0), "cart2sph launch");
    double *out = alloc_out(m);
    check(hipMemcpy(out, dsph.p, sizeof(double) * m, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return out;
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
        Int1eInput in = make_int1e_input(atm, atm_l, bas, bas_l, env, env_l);
        const size_t nn = (size_t)in.nao * in.nao;
        if (coord == 1) {
            DevMem d(nn);
            int1e_gpu_device(in, typec == 0 ? d.p : nullptr, typec == 1 ? d.p : nullptr, typec == 2 ? d.p : nullptr, 0);
            return to_spherical(2, bas, bas_l, d.p, in.nao);
        }
        double *R = alloc_out(nn);
        int1e_gpu(in, typec == 0 ? R : nullptr, typec == 1 ? R : nullptr, typec == 2 ? R : nullptr);
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
        if (coord == 1) {
            DevMem d(n * n * n * n);
            ctx.tensor_device(d.p);
            return to_spherical(4, bas, bas_l, d.p, (int)n);
        }
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
    Int1eInput in = make_int1e_input(atm, atm_l, bas, bas_l, env, env_l);
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
        Int1eInput in = make_int1e_input(atm, atm_l, bas, bas_l, env, env_l);
        std::vector<double> H(n * n), V(n * n);
        int1e_gpu(in, nullptr, H.data(), V.data());
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

}  // extern "C"

namespace {

// env split of the differentiable parameters (reference src/utils.rs:152-175):
// s1 = smallest exponent offset; s2 = coefficient offset + nprim of the
// shell that last raised the running maximum (same update rule).
void ref_split(const int *bas, size_t bas_l, size_t &s1, size_t &s2)
{
    long long mn = -1, mx = -1;
    for (size_t b = 0; b + 8 <= bas_l; b += 8) {
        const long long ngto = bas[b + 2], ex = bas[b + 5], cont = bas[b + 6];
        if (mn == -1 || mn > ex) mn = ex;
        if (mx == -1) mx = cont + ngto;
        else if (mx < cont) mx = cont + ngto;
    }
    s1 = mn < 0 ? 0 : (size_t)mn;
    s2 = mx < 0 ? 0 : (size_t)mx;
}

// Parts of the basis-parameter gradient of E = tr(P H) + 1/2 tr(P G[P]) at a
// converged density (reference src/dscf.rs:320-368, 517-537), each over the
// full env (entries outside the basis parameters stay zero), plus the
// nuclear-coordinate parts of the same terms.
struct GradParts {
    std::vector<double> dH, dR, dS;     // env-length
    std::vector<double> gH, gR, gS, gN; // 3*natm: core H, 2e, overlap (Pulay), nuclear repulsion
};

enum { GP_H = 1, GP_R = 2, GP_S = 4 };

GradParts grad_parts(Context *ctx, const Int1eInput &one, const double *P, size_t env_l, int what)
{
    GradParts g;
    const size_t na = one.charges.size();
    const int n = one.nao;
    const size_t nn = (size_t)n * n;
    // the one-electron derivative passes run together at the end, with one
    // staging of the basis on the GPU
    std::vector<Int1eGradPass> passes;
    std::vector<double> Q;
    if (what & GP_H) {
        g.dH.assign(env_l, 0.0);
        g.gH.assign(3 * na, 0.0);
        passes.push_back(Int1eGradPass{1, P, g.gH.data(), g.dH.data()});
        passes.push_back(Int1eGradPass{2, P, g.gH.data(), g.dH.data()});
    }
    if (what & GP_R) {
        g.dR.assign(env_l, 0.0);
        g.gR.assign(3 * na, 0.0);
        ctx->grad2e(P, g.gR.data(), g.dR.data(), env_l);
    }
    if (what & GP_S) {
        // Q = P F P with F = T + V + G[P]
        std::vector<double> F(nn), T(nn), G(nn);
        Q.assign(nn, 0.0);
        int1e_gpu(one, nullptr, T.data(), F.data());
        ctx->fock_host(P, G.data());
        for (size_t k = 0; k < nn; k++) F[k] += T[k] + G[k];
        pfp_gpu(n, P, F.data(), Q.data());
        // symmetrise against rounding so that the bra-derivative identity holds
        for (int a = 0; a < n; a++)
            for (int b = 0; b < a; b++) {
                const double v = 0.5 * (Q[(size_t)a * n + b] + Q[(size_t)b * n + a]);
                Q[(size_t)a * n + b] = Q[(size_t)b * n + a] = v;
            }
        g.dS.assign(env_l, 0.0);
        g.gS.assign(3 * na, 0.0);
        passes.push_back(Int1eGradPass{0, Q.data(), g.gS.data(), g.dS.data()});
    }
    int1e_grad_passes(one, passes, env_l);
    g.gN.assign(3 * na, 0.0);
    nuclear_repulsion_grad(one, g.gN.data());
    return g;
}

// One reference-ABI gradient call: build the context (no quartet screening,
// as the reference differentiates every canonical quartet), check P, combine
// the parts with the given weights and return the [s1, s2) slice.
double *grad_entry(const char *name, int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l,
                   double *P, size_t P_l, int what, double wH, double wR, double wS)
{
    try {
        if (!atm || !bas || !env || !P) {
            set_err(std::string(name) + ": null argument");
            return nullptr;
        }
        Int1eInput one = make_int1e_input(atm, atm_l, bas, bas_l, env, env_l);
        const size_t n = (size_t)one.nao;
        if (P_l != n * n) {
            set_err(std::string(name) + ": P has length " + std::to_string(P_l) + ", expected " + std::to_string(n * n));
            return nullptr;
        }
        // the GPU context is needed for the two-electron parts only
        std::unique_ptr<Context> ctx;
        if (what & (GP_R | GP_S)) {
            ctx.reset(new Context(atm, atm_l, bas, bas_l, env, env_l));
            ctx->set_schwarz(0.0);
        }
        GradParts g = grad_parts(ctx.get(), one, P, env_l, what);
        size_t s1, s2;
        ref_split(bas, bas_l, s1, s2);
        if (s2 < s1 || s2 > env_l) {
            set_err(std::string(name) + ": basis parameters lie outside env");
            return nullptr;
        }
        double *out = alloc_out(s2 - s1);
        for (size_t k = s1; k < s2; k++) {
            double v = 0.0;
            if (what & GP_H) v += wH * g.dH[k];
            if (what & GP_R) v += wR * g.dR[k];
            if (what & GP_S) v += wS * g.dS[k];
            out[k - s1] = v;
        }
        return out;
    } catch (std::exception &e) {
        set_err(std::string(name) + ": " + e.what());
        return nullptr;
    }
}

}  // namespace

extern "C" {

double *dS_u(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l)
{
    try {
        if (!atm || !bas || !env) {
            set_err("dS_u: null argument");
            return nullptr;
        }
        Int1eInput one = make_int1e_input(atm, atm_l, bas, bas_l, env, env_l);
        size_t s1, s2;
        ref_split(bas, bas_l, s1, s2);
        if (s2 < s1 || s2 > env_l) {
            set_err("dS_u: basis parameters lie outside env");
            return nullptr;
        }
        const size_t n = (size_t)one.nao, L = s2 - s1;
        double *out = alloc_out(n * n * L);
        overlap_param_derivs(one, s1, L, out);
        return out;
    } catch (std::exception &e) {
        set_err(std::string("dS_u: ") + e.what());
        return nullptr;
    }
}

double *grad_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l)
{
    return grad_entry("grad_c", atm, atm_l, bas, bas_l, env, env_l, P, P_l, GP_H | GP_R, 1.0, 1.0, 0.0);
}

double *dS_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l)
{
    return grad_entry("dS_c", atm, atm_l, bas, bas_l, env, env_l, P, P_l, GP_S, 0.0, 0.0, 1.0);
}

double *dHcore_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l)
{
    return grad_entry("dHcore_c", atm, atm_l, bas, bas_l, env, env_l, P, P_l, GP_H, 1.0, 0.0, 0.0);
}

double *dR_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l)
{
    return grad_entry("dR_c", atm, atm_l, bas, bas_l, env, env_l, P, P_l, GP_R, 0.0, 1.0, 0.0);
}

double *danalytical_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P,
                      size_t P_l)
{
    return grad_entry("danalytical_c", atm, atm_l, bas, bas_l, env, env_l, P, P_l, GP_H | GP_R | GP_S, 1.0, 1.0,
                      -0.5);
}

double *denergy_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l)
{
    return grad_entry("denergy_c", atm, atm_l, bas, bas_l, env, env_l, P, P_l, GP_H | GP_R | GP_S, 1.0, 1.0, -0.5);
}

double *grad_nuc_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l)
{
    try {
        if (!atm || !bas || !env || !P) {
            set_err("grad_nuc_c: null argument");
            return nullptr;
        }
        Int1eInput one = make_int1e_input(atm, atm_l, bas, bas_l, env, env_l);
        const size_t n = (size_t)one.nao;
        if (P_l != n * n) {
            set_err("grad_nuc_c: P has the wrong length");
            return nullptr;
        }
        Context ctx(atm, atm_l, bas, bas_l, env, env_l);
        ctx.set_schwarz(0.0);
        GradParts g = grad_parts(&ctx, one, P, env_l, GP_H | GP_R | GP_S);
        const size_t m = g.gN.size();
        double *out = alloc_out(m);
        for (size_t k = 0; k < m; k++) out[k] = g.gH[k] + g.gR[k] - 0.5 * g.gS[k] + g.gN[k];
        return out;
    } catch (std::exception &e) {
        set_err(std::string("grad_nuc_c: ") + e.what());
        return nullptr;
    }
}

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
        if (coord == 0) {
            if (P_l != n * n) {
                set_err("P has the wrong length (nao*nao cartesian)");
                return nullptr;
            }
            double *G = alloc_out(n * n);
            ctx.fock_host(P, G);
            return G;
        }
        // spherical (reference integral2e_fock with coord = 1): with C the
        // sph x cart rows, the spherical integrals are C x C x C x C of the
        // cartesian ones, so G_sph = C G_cart[C^T P C] C^T
        int ncart = 0;
        const std::vector<C2SRow> rows = c2s_rows(bas, bas_l, &ncart);
        const size_t ns = rows.size();
        if ((size_t)ncart != n || P_l != ns * ns) {
            set_err("P has the wrong length (nsph*nsph spherical)");
            return nullptr;
        }
        std::vector<double> T(ns * n, 0.0), Pc(n * n, 0.0), Gc(n * n);
        for (size_t mu = 0; mu < ns; mu++)  // T = P C  (ns x n)
            for (size_t nu = 0; nu < ns; nu++) {
                const double p = P[mu * ns + nu];
                if (p == 0.0) continue;
                for (int t = 0; t < rows[nu].n; t++) T[mu * n + rows[nu].cart[t]] += p * rows[nu].c[t];
            }
        for (size_t mu = 0; mu < ns; mu++)  // Pc = C^T T
            for (int t = 0; t < rows[mu].n; t++) {
                const size_t a = (size_t)rows[mu].cart[t];
                const double c = rows[mu].c[t];
                for (size_t b = 0; b < n; b++) Pc[a * n + b] += c * T[mu * n + b];
            }
        ctx.fock_host(Pc.data(), Gc.data());
        std::vector<double> U(ns * n, 0.0);  // U = C Gc  (ns x n)
        for (size_t mu = 0; mu < ns; mu++)
            for (int t = 0; t < rows[mu].n; t++) {
                const size_t a = (size_t)rows[mu].cart[t];
                const double c = rows[mu].c[t];
                for (size_t b = 0; b < n; b++) U[mu * n + b] += c * Gc[a * n + b];
            }
        double *G = alloc_out(ns * ns);  // G = U C^T
        for (size_t mu = 0; mu < ns; mu++)
            for (size_t nu = 0; nu < ns; nu++) {
                double v = 0.0;
                for (int t = 0; t < rows[nu].n; t++) v += U[mu * n + rows[nu].cart[t]] * rows[nu].c[t];
                G[mu * ns + nu] = v;
            }
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
        c->one = make_int1e_input(atm, atm_l, bas, bas_l, env, env_l);
        return c;
    } catch (std::exception &e) {
        set_err(e.what());
        return nullptr;
    }
}

void librint_ctx_destroy(librint_ctx *ctx) { delete ctx; }

int librint_ctx_nao(const librint_ctx *ctx) { return ctx ? ctx->ctx->nao() : -1; }

int librint_ctx_set_shard(librint_ctx *ctx, int nshards, int shard)
{
    try {
        if (!ctx) {
            set_err("null context");
            return -1;
        }
        ctx->ctx->set_shard(nshards, shard);
        return 0;
    } catch (std::exception &e) {
        set_err(e.what());
        return -1;
    }
}

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
        (void)dmax;  // kept for the ABI; the split follows each build's density blocks
        ctx->ctx->set_precision(mode, tau_mix, tau_skip);
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

}  // extern "C"

// Device scratch of librint_fock_device_noalloc: a file-scope array each GPU
// holds from the library's load until the process ends.  In host code its
// name is only a stand-in: the device address comes from
// hipGetSymbolAddress, once per device (with that device current), and is
// the only handle host code passes on.
namespace {
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)8 << 20;  // 64 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];
std::mutex g_scratch_mu;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

double *device_scratch()
{
    int dev = 0;
    librint::check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) throw std::runtime_error("device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        librint::check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = (double *)p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int librint_fock_device_noalloc(librint_ctx *ctx, const double *P_dev, double *G_dev, void *stream)
{
    try {
        if (!ctx || !P_dev || !G_dev) {
            set_err("null argument");
            return -1;
        }
        const size_t need = ctx->ctx->scratch_doubles_needed();
        if (need > NOALLOC_SCRATCH_DOUBLES) {
            set_err("librint_fock_device_noalloc: this build needs " + std::to_string(need * sizeof(double)) +
                    " bytes of scratch, above the " + std::to_string(NOALLOC_SCRATCH_DOUBLES * sizeof(double)) +
                    "-byte static buffer (use librint_fock_device)");
            return -1;
        }
        ctx->ctx->fock_device_scratch(P_dev, G_dev, (hipStream_t)stream, device_scratch(), NOALLOC_SCRATCH_DOUBLES);
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
    const double v[11] = {(double)s.quartets, (double)s.prim_quartets, (double)s.integrals, s.flops,
                          (double)s.candidates, (double)s.launches, (double)s.last_kernel_ms, ctx->ctx->schwarz(),
                          (double)s.last_grad_ms, (double)s.scf_purified, (double)s.streams};
    int k = 0;
    for (; k < n && k < 11; k++) out[k] = v[k];
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
        const double v[14] = {(double)c.la,        (double)c.lb,    (double)c.lc,      (double)c.ld,
                              (double)c.quartets,  (double)c.prim_quartets, c.flops,  (double)c.lanes,
                              (double)c.ms,        (double)c.count, (double)c.fp32,    (double)c.variant,
                              (double)c.drow,      (double)c.reduce_ms};
        for (int j = 0; j < 14; j++) out[14 * k + j] = v[j];
    }
    return (int)L.size();
}

int librint_ctx_gradient(librint_ctx *ctx, const double *P, double *gnuc, double *denv, double *parts)
{
    try {
        if (!ctx || !P) {
            set_err("null argument");
            return -1;
        }
        Context &c = *ctx->ctx;
        const size_t env_l = c.env().size();
        GradParts g = grad_parts(&c, ctx->one, P, env_l, GP_H | GP_R | GP_S);
        const size_t m = g.gN.size();
        if (gnuc)
            for (size_t k = 0; k < m; k++) gnuc[k] = g.gH[k] + g.gR[k] - 0.5 * g.gS[k] + g.gN[k];
        if (denv)
            for (size_t k = 0; k < env_l; k++) denv[k] = g.dH[k] + g.dR[k] - 0.5 * g.dS[k];
        if (parts) {
            // [gH | gR | gS | gN] (each 3*natm) then [dH | dR | dS] (each env_l)
            double *o = parts;
            for (const auto *v : {&g.gH, &g.gR, &g.gS, &g.gN, &g.dH, &g.dR, &g.dS}) {
                std::memcpy(o, v->data(), sizeof(double) * v->size());
                o += v->size();
            }
        }
        return 0;
    } catch (std::exception &e) {
        set_err(e.what());
        return -1;
    }
}

double librint_scf_ex(librint_ctx *ctx, int nelec, int imax, double conv, double conv_e, int guess, int diis,
                      double *P_inout, int *iters)
{
    try {
        if (!ctx) {
            set_err("null context");
            return NAN;
        }
        ScfOptions o;
        o.guess = guess;
        o.diis = diis;
        o.conv_e = conv_e;
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

static int mlp_status(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    set_err(std::string(what) + ": " + hipGetErrorString(e));
    return -1;
}

int librint_mlp_gemm(int M, int N, int K, const uint16_t *A, int lda, int transA, const uint16_t *B, int ldb,
                     int transB, float *C, int ldc, const float *bias, int act, int accumulate, uint16_t *Cb, int ldcb,
                     void *stream)
{
    return mlp_status(launch_gemm_bf16(M, N, K, A, lda, transA, B, ldb, transB, C, ldc, bias, act, accumulate, Cb, ldcb,
                                       (hipStream_t)stream),
                      "librint_mlp_gemm");
}

int librint_mlp_relu_bwd(size_t n, const float *dz, const float *h, float *dpre, uint16_t *dpre_b, void *stream)
{
    return mlp_status(launch_relu_bwd(n, dz, h, dpre, dpre_b, (hipStream_t)stream), "librint_mlp_relu_bwd");
}

int librint_mlp_colsum(int rows, int cols, const float *x, int ld, float *out, int accumulate, void *stream)
{
    return mlp_status(launch_colsum(rows, cols, x, ld, out, accumulate, (hipStream_t)stream), "librint_mlp_colsum");
}

int librint_mlp_mse(int rows, int cols, const float *y, int ldy, const float *t, const float *mask, int ldt,
                    float scale, float *dy, uint16_t *dy_b, double *loss, void *stream)
{
    return mlp_status(launch_mse(rows, cols, y, ldy, t, mask, ldt, scale, dy, dy_b, loss, (hipStream_t)stream),
                      "librint_mlp_mse");
}

int librint_mlp_adam(size_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                     int step, uint16_t *wb, void *stream)
{
    return mlp_status(launch_adam(n, w, g, m, v, lr, b1, b2, eps, step, wb, (hipStream_t)stream), "librint_mlp_adam");
}

int librint_mlp_to_bf16(size_t n, const float *x, uint16_t *y, void *stream)
{
    return mlp_status(launch_f32_to_bf16(n, x, y, (hipStream_t)stream), "librint_mlp_to_bf16");
}

}  // extern "C"
