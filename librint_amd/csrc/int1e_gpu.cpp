// Host driver of the one-electron kernels (int1e_kernel.hip): stages shells, env and
// point charges on the current GPU, launches one kernel per shell-pair class
// and copies the matrices back.
#include "int1e_gpu.h"

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

}  // namespace

void int1e_gpu_device(const Int1eInput &in, double *dS, double *dT, double *dV, hipStream_t st)
{
    const int ns = (int)in.shells.size();
    if (!dS && !dT && !dV) return;
    std::vector<Shell1Dev> sh(ns);
    for (int s = 0; s < ns; s++) {
        const Shell1 &x = in.shells[s];
        if (x.l < 0 || x.l > 4) throw std::invalid_argument("one-electron integrals on the GPU cover l <= 4");
        sh[s] = Shell1Dev{x.l, x.nprim, x.pexp, x.pcoef, x.ao, 0, {x.r[0], x.r[1], x.r[2]}};
    }
    std::vector<double4> nuc(in.charges.size());
    for (size_t c = 0; c < nuc.size(); c++)
        nuc[c] = make_double4(in.centers[c][0], in.centers[c][1], in.centers[c][2], in.charges[c]);
    // ordered shell pairs, grouped by class
    std::vector<int2> pairs;
    int first[5][5], count[5][5];
    for (int li = 0; li < 5; li++)
        for (int lj = 0; lj < 5; lj++) {
            first[li][lj] = (int)pairs.size();
            for (int i = 0; i < ns; i++)
                if (sh[i].l == li)
                    for (int j = 0; j < ns; j++)
                        if (sh[j].l == lj) pairs.push_back(make_int2(i, j));
            count[li][lj] = (int)pairs.size() - first[li][lj];
        }
    DevBuf<Shell1Dev> dsh(sh);
    DevBuf<double> denv(in.env);
    DevBuf<double4> dnuc(nuc);
    DevBuf<int2> dpairs(pairs);
    for (int li = 0; li < 5; li++)
        for (int lj = 0; lj < 5; lj++)
            ck(launch_int1e(li, lj, dpairs.p + first[li][lj], count[li][lj], dsh.p, denv.p, dnuc.p, (int)nuc.size(),
                            in.nao, dS, dT, dV, st),
               "int1e launch");
    // the staging buffers are freed on return: finish the launches first
    ck(hipStreamSynchronize(st), "hipStreamSynchronize");
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
