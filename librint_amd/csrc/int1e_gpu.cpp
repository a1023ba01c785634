// Host driver of the one-electron kernels (int1e_kernel.hip): stages shells, env and
// point charges on the current GPU, launches one kernel per shell-pair class
// and copies the matrices (or the derivative sums) back.
#include "int1e_gpu.h"

#include <memory>
#include <stdexcept>
#include <string>

namespace librint {

namespace {

void ck(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    explicit DevBuf(size_t n)
    {
        if (n) ck(hipMalloc(&p, sizeof(T) * n), "hipMalloc");
    }
    DevBuf(const std::vector<T> &v) : DevBuf(v.size())
    {
        if (!v.empty()) ck(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

// Shells (atom index in pad), point charges and ordered shell pairs grouped
// by (li, lj) for lmax.
struct Staged {
    std::vector<Shell1Dev> sh;
    std::vector<double4> nuc;
    std::vector<int2> pairs;
    int first[5][5] = {}, count[5][5] = {};
};

Staged stage(const Int1eInput &in, int lmax, const char *what)
{
    Staged g;
    const int ns = (int)in.shells.size();
    g.sh.resize(ns);
    for (int s = 0; s < ns; s++) {
        const Shell1 &x = in.shells[s];
        if (x.l < 0 || x.l > lmax) throw std::invalid_argument(what);
        g.sh[s] = Shell1Dev{x.l, x.nprim, x.pexp, x.pcoef, x.ao, x.atom, {x.r[0], x.r[1], x.r[2]}};
    }
    g.nuc.resize(in.charges.size());
    for (size_t c = 0; c < g.nuc.size(); c++)
        g.nuc[c] = make_double4(in.centers[c][0], in.centers[c][1], in.centers[c][2], in.charges[c]);
    for (int li = 0; li <= lmax; li++)
        for (int lj = 0; lj <= lmax; lj++) {
            g.first[li][lj] = (int)g.pairs.size();
            for (int i = 0; i < ns; i++)
                if (g.sh[i].l == li)
                    for (int j = 0; j < ns; j++)
                        if (g.sh[j].l == lj) g.pairs.push_back(make_int2(i, j));
            g.count[li][lj] = (int)g.pairs.size() - g.first[li][lj];
        }
    return g;
}

}  // namespace

void int1e_gpu_device(const Int1eInput &in, double *dS, double *dT, double *dV, hipStream_t st)
{
    if (!dS && !dT && !dV) return;
    const Staged g = stage(in, 4, "one-electron integrals on the GPU cover l <= 4");
    DevBuf<Shell1Dev> dsh(g.sh);
    DevBuf<double> denv(in.env);
    DevBuf<double4> dnuc(g.nuc);
    DevBuf<int2> dpairs(g.pairs);
    for (int li = 0; li < 5; li++)
        for (int lj = 0; lj < 5; lj++)
            ck(launch_int1e(li, lj, dpairs.p + g.first[li][lj], g.count[li][lj], dsh.p, denv.p, dnuc.p,
                            (int)g.nuc.size(), in.nao, dS, dT, dV, st),
               "int1e launch");
    // the staging buffers are freed on return: finish the launches first
    ck(hipStreamSynchronize(st), "hipStreamSynchronize");
}

// Derivative passes on the GPU with one staging of the shells, env and
// point charges: each pass adds its W-weighted sums into gatom/denv (host
// arrays; passes with the same pair of targets share one device
// accumulator, copied back and added once), or writes the dS matrix (typ 0,
// W null).
static void deriv_gpu(const Int1eInput &in, const std::vector<Int1eGradPass> &passes, size_t env_l, double *dSout,
                      size_t s1, size_t L)
{
    const Staged g = stage(in, 4, "one-electron derivatives cover l <= 4");
    const size_t nn = (size_t)in.nao * in.nao;
    const size_t na = in.charges.size();
    DevBuf<Shell1Dev> dsh(g.sh);
    DevBuf<double> denv_in(in.env);
    DevBuf<double4> dnuc(g.nuc);
    DevBuf<int2> dpairs(g.pairs);
    bool anyW = false;
    for (const Int1eGradPass &q : passes) anyW = anyW || q.W;
    DevBuf<double> dW(anyW ? nn : 0), ddS(dSout ? nn * L : 0);
    if (dSout && nn * L) ck(hipMemset(ddS.p, 0, sizeof(double) * nn * L), "hipMemset");
    // one device accumulator per distinct (gatom, denv) target pair
    struct Acc {
        double *gatom, *denv;
        std::unique_ptr<DevBuf<double>> dga, dde;
    };
    std::vector<Acc> acc;
    auto acc_of = [&](double *gatom, double *denv) -> Acc & {
        for (Acc &a : acc)
            if (a.gatom == gatom && a.denv == denv) return a;
        acc.push_back(Acc{gatom, denv, std::make_unique<DevBuf<double>>(gatom ? 3 * na : 0),
                          std::make_unique<DevBuf<double>>(denv ? env_l : 0)});
        Acc &a = acc.back();
        if (gatom && na) ck(hipMemset(a.dga->p, 0, sizeof(double) * 3 * na), "hipMemset");
        if (denv && env_l) ck(hipMemset(a.dde->p, 0, sizeof(double) * env_l), "hipMemset");
        return a;
    };
    for (const Int1eGradPass &q : passes) {
        // the null stream orders this copy after the previous pass's kernels
        if (q.W) ck(hipMemcpy(dW.p, q.W, sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy H2D");
        Acc &a = acc_of(q.gatom, q.denv);
        for (int li = 0; li < 5; li++)
            for (int lj = 0; lj < 5; lj++)
                ck(launch_int1e_deriv(li, lj, dpairs.p + g.first[li][lj], g.count[li][lj], dsh.p, denv_in.p, dnuc.p,
                                      (int)g.nuc.size(), in.nao, q.typ, q.W ? dW.p : nullptr, a.dga->p, a.dde->p,
                                      (long long)env_l, q.W ? nullptr : ddS.p, (long long)s1, (long long)L, 0),
                   "one-electron derivative launch");
    }
    for (Acc &a : acc) {
        if (a.gatom && na) {
            std::vector<double> h(3 * na);
            ck(hipMemcpy(h.data(), a.dga->p, sizeof(double) * 3 * na, hipMemcpyDeviceToHost), "hipMemcpy D2H");
            for (size_t k = 0; k < 3 * na; k++) a.gatom[k] += h[k];
        }
        if (a.denv && env_l) {
            std::vector<double> h(env_l);
            ck(hipMemcpy(h.data(), a.dde->p, sizeof(double) * env_l, hipMemcpyDeviceToHost), "hipMemcpy D2H");
            for (size_t k = 0; k < env_l; k++) a.denv[k] += h[k];
        }
    }
    if (dSout && nn * L) ck(hipMemcpy(dSout, ddS.p, sizeof(double) * nn * L, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

void int1e_grad_passes(const Int1eInput &in, const std::vector<Int1eGradPass> &passes, size_t env_l)
{
    for (const Int1eGradPass &q : passes) {
        if (q.typ < 0 || q.typ > 2) throw std::invalid_argument("one-electron derivative type must be 0, 1 or 2");
        if (!q.W) throw std::invalid_argument("null weight matrix");
    }
    if (!passes.empty()) deriv_gpu(in, passes, env_l, nullptr, 0, 0);
}

void int1e_grad(const Int1eInput &in, int typ, const double *W, double *gatom, double *denv, size_t env_l)
{
    int1e_grad_passes(in, {Int1eGradPass{typ, W, gatom, denv}}, env_l);
}

void overlap_param_derivs(const Int1eInput &in, size_t s1, size_t L, double *dS)
{
    deriv_gpu(in, {Int1eGradPass{0, nullptr, nullptr, nullptr}}, 0, dS, s1, L);
}

void int1e_gpu(const Int1eInput &in, double *S, double *T, double *V)
{
    const size_t nn = (size_t)in.nao * in.nao;
    if (!S && !T && !V) return;
    DevBuf<double> dS(S ? nn : 0), dT(T ? nn : 0), dV(V ? nn : 0);
    int1e_gpu_device(in, dS.p, dT.p, dV.p, 0);
    if (S) ck(hipMemcpy(S, dS.p, sizeof(double) * nn, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    if (T) ck(hipMemcpy(T, dT.p, sizeof(double) * nn, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    if (V) ck(hipMemcpy(V, dV.p, sizeof(double) * nn, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

}  // namespace librint
