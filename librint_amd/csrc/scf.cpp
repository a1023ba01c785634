// RHF driver on the GPU: one-electron matrices from the GPU kernels, the
// fused J/K Fock build, and the dense linear algebra (orthogonaliser,
// F' = X^T F X, eigenvectors, density, DIIS error vectors) on rocBLAS and
// rocSOLVER; only the small DIIS system is solved on the host.
//
// Reference mode (default, reference src/scf.rs:743-799): P0 = identity,
// plain Roothaan iteration F -> X^T F X -> eigh -> aufbau P, stop when
// |P - P_old|_F / 2 <= conv, then the validity checks of src/scf.rs:572-600
// and the energy E = tr(P(T+V)) + 1/2 tr(P G[P]) + E_nuc (src/scf.rs:827-956).
// Options: core-Hamiltonian or caller-supplied starting density, Pulay DIIS
// extrapolation and an energy criterion (not in the reference; off unless
// requested).
//
// Storage is column-major (rocBLAS); every matrix that crosses the Fock build
// or the caller is symmetric, so the layouts agree there.
#include "scf.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>

#include "int1e_gpu.h"

namespace librint {

namespace {

void hck(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void bck(rocblas_status s, const char *what)
{
    if (s != rocblas_status_success) throw std::runtime_error(std::string(what) + ": rocBLAS status " + std::to_string((int)s));
}

struct Dev {
    double *p = nullptr;
    explicit Dev(size_t n) { hck(hipMalloc(&p, sizeof(double) * std::max<size_t>(n, 1)), "hipMalloc"); }
    ~Dev()
    {
        if (p) (void)hipFree(p);
    }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
};

struct Blas {
    rocblas_handle h = nullptr;
    Blas()
    {
        bck(rocblas_create_handle(&h), "rocblas_create_handle");
        bck(rocblas_set_stream(h, 0), "rocblas_set_stream");
    }
    ~Blas()
    {
        if (h) rocblas_destroy_handle(h);
    }
    // C = alpha op(A) op(B) + beta C, all n x n
    void gemm(int n, bool ta, bool tb, double alpha, const double *A, const double *B, double beta, double *C,
              int k = -1)
    {
        const int kk = k < 0 ? n : k;
        bck(rocblas_dgemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                          tb ? rocblas_operation_transpose : rocblas_operation_none, n, n, kk, &alpha, A, n, B, n,
                          &beta, C, n),
            "rocblas_dgemm");
    }
    double dot(size_t m, const double *x, const double *y)
    {
        double r = 0.0;
        bck(rocblas_ddot(h, (rocblas_int)m, x, 1, y, 1, &r), "rocblas_ddot");
        return r;
    }
    void axpy(size_t m, double a, const double *x, double *y)
    {
        bck(rocblas_daxpy(h, (rocblas_int)m, &a, x, 1, y, 1), "rocblas_daxpy");
    }
    double nrm2(size_t m, const double *x)
    {
        double r = 0.0;
        bck(rocblas_dnrm2(h, (rocblas_int)m, x, 1, &r), "rocblas_dnrm2");
        return r;
    }
    double amax(size_t m, const double *x)
    {
        rocblas_int i = 0;
        bck(rocblas_idamax(h, (rocblas_int)m, x, 1, &i), "rocblas_idamax");
        double v = 0.0;
        if (i >= 1) hck(hipMemcpy(&v, x + (i - 1), sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        return std::fabs(v);
    }
    void copy(size_t m, const double *x, double *y)
    {
        hck(hipMemcpyAsync(y, x, sizeof(double) * m, hipMemcpyDeviceToDevice, 0), "hipMemcpyAsync");
    }
    // eigenvalues ascending into w (device), eigenvectors over A (columns)
    void eigh(int n, double *A, double *w, double *e, rocblas_int *info)
    {
        bck(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, A, n, w, e, info), "rocsolver_dsyevd");
        rocblas_int hinfo = 0;
        hck(hipMemcpy(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        if (hinfo != 0) throw std::runtime_error("rocsolver_dsyevd did not converge (info " + std::to_string(hinfo) + ")");
    }
};

__global__ void add_kernel(size_t m, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) c[i] = a[i] + b[i];
}

// X[:, k] = U[:, k] / sqrt(s[k])
__global__ void scale_cols_kernel(int n, const double *__restrict__ U, const double *__restrict__ s,
                                  double *__restrict__ X)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n) return;
    const int k = (int)(i / n);
    X[i] = U[i] * rsqrt(s[k]);
}

// e = A - A^T
__global__ void antisym_kernel(int n, const double *__restrict__ A, double *__restrict__ e)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n) return;
    const int r = (int)(i % n), c = (int)(i / n);
    e[i] = A[i] - A[(size_t)r * n + c];
}

// Gershgorin bounds of an n x n matrix (one workgroup): out[0] =
// min_j (A_jj - r_j), out[1] = max_j (A_jj + r_j), r_j the off-diagonal
// absolute sum of row j (the row discs contain the spectrum as the column
// discs do).  Thread j walks row j of the column-major matrix, so the
// lanes of a wavefront read consecutive elements of each column.
__global__ __launch_bounds__(1024) void gersh_kernel(int n, const double *__restrict__ A, double *__restrict__ out)
{
    __shared__ double lo[1024], hi[1024];
    double l = 1e300, h = -1e300;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double r = 0.0;
        for (int i = 0; i < n; i++) r += fabs(A[(size_t)i * n + j]);
        const double d = A[(size_t)j * n + j];
        r -= fabs(d);
        l = fmin(l, d - r);
        h = fmax(h, d + r);
    }
    lo[threadIdx.x] = l;
    hi[threadIdx.x] = h;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + s]);
            hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = lo[0];
        out[1] = hi[0];
    }
}

// X = (emax I - A) / (emax - emin): the spectrum mapped into [0, 1], the
// lowest eigenvalues of A nearest 1
__global__ void sp2_init_kernel(int n, const double *__restrict__ A, double emin, double emax, double *__restrict__ X)
{
    const size_t nn = (size_t)n * n;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const size_t i = k % n, j = k / n;
    X[k] = ((i == j ? emax : 0.0) - A[k]) / (emax - emin);
}

// State of the purification on the device: the idempotency errors of the
// last two steps, the trace at the stop, the steps taken, the last step's
// choice and whether it has stopped.
struct Sp2State {
    double hist0, hist1;
    double tr;
    int steps, choice, done, pad;
};

// One purification step's decision (one workgroup), nothing once stopped:
// t1 = tr X, t2 = tr X2 (X2 = X^2); stop when the idempotency error t1 - t2
// has reached rounding level, where it no longer falls over two steps (the
// stopping rule of Kruchinina, Rudberg and Rubensson, J. Chem. Theory
// Comput. 12, 5788 (2016)), and is below idem_floor, so that every
// eigenvalue of the projector is within about 1e-11 of 0 or 1 (a looser stop
// left the SCF density unable to settle below |dP| ~ 1e-9); else choose X^2
// (0) or 2X - X^2 (1), whichever brings the trace nearer to nocc.
__global__ __launch_bounds__(1024) void sp2_step_kernel(int n, const double *__restrict__ X,
                                                        const double *__restrict__ X2, Sp2State *st, double nocc,
                                                        double idem_floor)
{
    if (st->done) return;
    __shared__ double a[1024], b[1024];
    double s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        s1 += X[(size_t)i * n + i];
        s2 += X2[(size_t)i * n + i];
    }
    a[threadIdx.x] = s1;
    b[threadIdx.x] = s2;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            a[threadIdx.x] += a[threadIdx.x + s];
            b[threadIdx.x] += b[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double t1 = a[0], t2 = b[0], idem = t1 - t2;
        if ((st->steps >= 2 && idem < idem_floor && idem >= st->hist1) || idem <= 0.0) {
            st->done = 1;
            st->tr = t1;
        } else {
            st->hist1 = st->hist0;
            st->hist0 = idem;
            st->choice = fabs(t2 - nocc) < fabs(2.0 * t1 - t2 - nocc) ? 0 : 1;
            st->steps++;
        }
    }
}

// X = X2 (choice 0) or 2 X - X2 (choice 1), nothing once stopped
__global__ void sp2_apply_kernel(size_t m, double *__restrict__ X, const double *__restrict__ X2,
                                 const Sp2State *st)
{
    if (st->done) return;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) X[k] = st->choice == 0 ? X2[k] : 2.0 * X[k] - X2[k];
}

__global__ void identity_kernel(int n, double *__restrict__ A)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n) return;
    A[i] = (i % n == i / n) ? 1.0 : 0.0;
}

unsigned blocks(size_t m) { return (unsigned)((m + 255) / 256); }

bool solve_linear(int n, std::vector<double> A, std::vector<double> b, std::vector<double> &x)
{
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(A[r * n + c]) > std::fabs(A[p * n + c])) p = r;
        if (std::fabs(A[p * n + c]) < 1e-300) return false;
        if (p != c) {
            for (int k = 0; k < n; k++) std::swap(A[c * n + k], A[p * n + k]);
            std::swap(b[c], b[p]);
        }
        for (int r = c + 1; r < n; r++) {
            const double f = A[r * n + c] / A[c * n + c];
            for (int k = c; k < n; k++) A[r * n + k] -= f * A[c * n + k];
            b[r] -= f * b[c];
        }
    }
    x.assign(n, 0.0);
    for (int r = n - 1; r >= 0; r--) {
        double s = b[r];
        for (int k = r + 1; k < n; k++) s -= A[r * n + k] * x[k];
        x[r] = s / A[r * n + r];
    }
    return true;
}

}  // namespace

// rocBLAS handle for the products outside the SCF loop: one per host thread
// and device, created on first use and kept (never destroyed) until the
// process ends, so that repeated calls do not pay for creating it
static Blas &thread_blas()
{
    thread_local std::map<int, Blas *> handles;
    int dev = 0;
    hck(hipGetDevice(&dev), "hipGetDevice");
    Blas *&b = handles[dev];
    if (!b) b = new Blas();
    return *b;
}

void pfp_gpu(int n, const double *P, const double *F, double *Q)
{
    const size_t nn = (size_t)n * n;
    Blas &B = thread_blas();
    Dev dP(nn), dF(nn), dT(nn), dQ(nn);
    hck(hipMemcpy(dP.p, P, sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy H2D");
    hck(hipMemcpy(dF.p, F, sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy H2D");
    B.gemm(n, false, false, 1.0, dP.p, dF.p, 0.0, dT.p);
    B.gemm(n, false, false, 1.0, dT.p, dP.p, 0.0, dQ.p);
    hck(hipMemcpy(Q, dQ.p, sizeof(double) * nn, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

double energy_fast(Context &ctx, const Int1eInput &in, const double *P, const double *H)
{
    const int n = ctx.nao();
    const size_t nn = (size_t)n * n;
    std::vector<double> G(nn);
    ctx.fock_host(P, G.data());
    double E = 0;
    for (size_t i = 0; i < nn; i++) E += P[i] * (H[i] + 0.5 * G[i]);
    return E + nuclear_repulsion(in);
}

ScfOut scf_run(Context &ctx, const Int1eInput &in, int nelec, int imax, double conv, const ScfOptions &o)
{
    ScfOut out;
    if (ctx.nshards() > 1) throw std::invalid_argument("the SCF needs the whole Fock build (nshards = 1)");
    const int n = ctx.nao();
    const size_t nn = (size_t)n * n;
    if (in.nao != n) throw std::runtime_error("basis mismatch between 1e and 2e inputs");
    if (nelec % 2) throw std::invalid_argument("closed-shell RHF needs an even electron count");
    if (nelec / 2 > n) throw std::invalid_argument("more occupied orbitals than basis functions");
    const int nocc = nelec / 2;
    Blas B;
    Dev S(nn), H(nn), V(nn), X(nn), T1(nn), T2(nn), F(nn), G(nn), P(nn), Pold(nn), w(n), e(n);
    rocblas_int *info = nullptr;
    hck(hipMalloc(&info, sizeof(rocblas_int)), "hipMalloc");
    struct InfoFree {
        rocblas_int *p;
        ~InfoFree() { (void)hipFree(p); }
    } info_free{info};
    int1e_gpu_device(in, S.p, H.p, V.p, 0);
    B.axpy(nn, 1.0, V.p, H.p);  // H = T + V

    // X = U s^{-1/2}
    B.copy(nn, S.p, T1.p);
    B.eigh(n, T1.p, w.p, e.p, info);
    double smin = 0.0;
    hck(hipMemcpy(&smin, w.p, sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    if (smin < 1e-8) {
        out.status = 3;
        out.msg = "overlap matrix is near-singular (smallest eigenvalue < 1e-8)";
        return out;
    }
    hipLaunchKernelGGL(scale_cols_kernel, dim3(blocks(nn)), dim3(256), 0, 0, n, T1.p, w.p, X.p);
    hck(hipGetLastError(), "scale_cols launch");

    // The occupied-space projector of F' = X^T F X by trace-correcting
    // purification (A. M. N. Niklasson, Phys. Rev. B 66, 155115 (2002)):
    // the spectrum mapped into [0, 1] by its Gershgorin bounds, then
    // X <- X^2 or 2X - X^2, whichever brings tr X nearer to nocc, until the
    // idempotency error tr X - tr X^2 stops falling.  Matrix products only
    // (rocBLAS dgemm), against the divide-and-conquer eigensolver's
    // tridiagonal reduction.  Returns false (the caller diagonalises) when it
    // does not settle on a projector of trace nocc: a closed or vanishing
    // gap, where the aufbau occupation is not a projector's limit.
    // The steps' decisions run on the device (sp2_step_kernel); the host
    // reads the state back every fourth step only, so a step costs its
    // product and two small kernels instead of a round trip, and at most
    // three products run past the stop (they change nothing).
    Dev SPa(nn), SPb(nn), sc2(2), sp2s((sizeof(Sp2State) + sizeof(double) - 1) / sizeof(double));
    Sp2State *dst = (Sp2State *)sp2s.p;
    auto purify = [&](const double *Fo, double *Proj) -> bool {
        double hb[2];
        hipLaunchKernelGGL(gersh_kernel, dim3(1), dim3(1024), 0, 0, n, Fo, sc2.p);
        hck(hipGetLastError(), "gersh launch");
        hck(hipMemcpy(hb, sc2.p, sizeof(hb), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        if (!(hb[1] > hb[0])) return false;
        hipLaunchKernelGGL(sp2_init_kernel, dim3(blocks(nn)), dim3(256), 0, 0, n, Fo, hb[0], hb[1], Proj);
        hck(hipGetLastError(), "sp2 init launch");
        const Sp2State init{1e300, 1e300, 0.0, 0, 0, 0, 0};
        hck(hipMemcpy(dst, &init, sizeof(init), hipMemcpyHostToDevice), "hipMemcpy H2D");
        const double idem_floor = std::max(1e-11, 1e-14 * n);
        for (int k = 0; k < 120; k++) {
            B.gemm(n, false, false, 1.0, Proj, Proj, 0.0, SPb.p);  // X^2
            hipLaunchKernelGGL(sp2_step_kernel, dim3(1), dim3(1024), 0, 0, n, Proj, SPb.p, dst, (double)nocc,
                               idem_floor);
            hck(hipGetLastError(), "sp2 step launch");
            hipLaunchKernelGGL(sp2_apply_kernel, dim3(blocks(nn)), dim3(256), 0, 0, nn, Proj, SPb.p, dst);
            hck(hipGetLastError(), "sp2 update launch");
            if ((k & 3) == 3 || k == 119) {
                Sp2State h;
                hck(hipMemcpy(&h, dst, sizeof(h), hipMemcpyDeviceToHost), "hipMemcpy D2H");
                if (h.done) return std::fabs(h.tr - nocc) < 1e-6;
            }
        }
        return false;
    };
    // LIBRINT_SCF_EIGH=1 diagonalises in every iteration of the DIIS mode too
    const bool use_purify = o.diis && o.guess != 0 && !getenv("LIBRINT_SCF_EIGH");
    int npurified = 0;

    // P from F: F' = X^T F X, eigenvectors C', C = X C', P = 2 C_occ C_occ^T
    // (or, in the DIIS mode while the density still changes by more than
    // 100 conv, P = 2 X Proj X^T with the purified projector: the projector
    // is idempotent to rounding level but not to the eigenvectors' 1e-15,
    // and X = U s^{-1/2} magnifies what is left, so the last cycles, which
    // decide the converged density, diagonalise)
    auto density_from_F = [&](const double *Fm, double dlast) {
        B.gemm(n, false, false, 1.0, Fm, X.p, 0.0, T1.p);   // F X
        B.gemm(n, true, false, 1.0, X.p, T1.p, 0.0, T2.p);  // X^T F X
        if (use_purify && dlast > 100.0 * conv && purify(T2.p, SPa.p)) {
            B.gemm(n, false, false, 1.0, X.p, SPa.p, 0.0, T1.p);  // X Proj
            B.gemm(n, false, true, 2.0, T1.p, X.p, 0.0, P.p);     // 2 X Proj X^T
            npurified++;
            return;
        }
        B.eigh(n, T2.p, w.p, e.p, info);
        B.gemm(n, false, false, 1.0, X.p, T2.p, 0.0, T1.p);  // C
        B.gemm(n, false, true, 2.0, T1.p, T1.p, 0.0, P.p, nocc);
    };
    if (o.guess == 0) {
        hipLaunchKernelGGL(identity_kernel, dim3(blocks(nn)), dim3(256), 0, 0, n, P.p);
        hck(hipGetLastError(), "identity launch");
    } else if (o.guess == 1) {
        density_from_F(H.p, 1.0);
    } else {
        if (!o.P0) throw std::invalid_argument("guess=2 needs a starting density");
        hck(hipMemcpy(P.p, o.P0, sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    const int NDIIS = 8;
    std::vector<Dev *> dF, dE;
    struct Hist {
        std::vector<Dev *> &a, &b;
        ~Hist()
        {
            for (Dev *d : a) delete d;
            for (Dev *d : b) delete d;
        }
    } hist{dF, dE};
    std::vector<double> Bm;  // DIIS overlaps, row-major NDIIS x NDIIS of the current history
    double delta = 1.0, Eprev = 0.0, dEn = 1.0;
    const double Enuc = nuclear_repulsion(in);
    int it = 0;
    auto converged = [&]() { return delta <= conv && (o.conv_e <= 0.0 || std::fabs(dEn) <= o.conv_e); };
    while (!converged() && it < imax) {
        B.copy(nn, P.p, Pold.p);
        ctx.fock_device(P.p, G.p, 0);
        hipLaunchKernelGGL(add_kernel, dim3(blocks(nn)), dim3(256), 0, 0, nn, H.p, G.p, F.p);
        hck(hipGetLastError(), "add launch");
        if (o.conv_e > 0.0) {
            const double Ek = Enuc + B.dot(nn, P.p, H.p) + 0.5 * B.dot(nn, P.p, G.p);
            dEn = it == 0 ? 1.0 : Ek - Eprev;
            Eprev = Ek;
        }
        if (o.diis) {
            // e = X^T (F P S - S P F) X
            B.gemm(n, false, false, 1.0, F.p, P.p, 0.0, T1.p);
            B.gemm(n, false, false, 1.0, T1.p, S.p, 0.0, T2.p);
            hipLaunchKernelGGL(antisym_kernel, dim3(blocks(nn)), dim3(256), 0, 0, n, T2.p, T1.p);
            hck(hipGetLastError(), "antisym launch");
            B.gemm(n, true, false, 1.0, X.p, T1.p, 0.0, T2.p);
            Dev *en = nullptr, *fn = nullptr;
            if ((int)dF.size() == NDIIS) {  // recycle the oldest buffers
                fn = dF.front();
                en = dE.front();
                dF.erase(dF.begin());
                dE.erase(dE.begin());
                const int m0 = NDIIS;
                std::vector<double> nb((NDIIS) * (NDIIS), 0.0);
                for (int i = 1; i < m0; i++)
                    for (int j = 1; j < m0; j++) nb[(i - 1) * NDIIS + (j - 1)] = Bm[i * NDIIS + j];
                Bm.swap(nb);
            } else {
                fn = new Dev(nn);
                en = new Dev(nn);
            }
            B.gemm(n, false, false, 1.0, T2.p, X.p, 0.0, en->p);  // X^T (FPS - SPF) X
            B.copy(nn, F.p, fn->p);
            dF.push_back(fn);
            dE.push_back(en);
            const int m = (int)dF.size();
            if ((int)Bm.size() != NDIIS * NDIIS) Bm.assign(NDIIS * NDIIS, 0.0);
            for (int j = 0; j < m; j++) {
                const double d = B.dot(nn, dE[m - 1]->p, dE[j]->p);
                Bm[(m - 1) * NDIIS + j] = Bm[j * NDIIS + (m - 1)] = d;
            }
            if (m >= 2) {
                std::vector<double> A((m + 1) * (m + 1), 0.0), rhs(m + 1, 0.0), c;
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++) A[i * (m + 1) + j] = Bm[i * NDIIS + j];
                for (int i = 0; i < m; i++) A[i * (m + 1) + m] = A[m * (m + 1) + i] = -1.0;
                rhs[m] = -1.0;
                if (solve_linear(m + 1, A, rhs, c)) {
                    hck(hipMemsetAsync(F.p, 0, sizeof(double) * nn, 0), "hipMemsetAsync");
                    for (int i = 0; i < m; i++) B.axpy(nn, c[i], dF[i]->p, F.p);
                }
            }
        }
        density_from_F(F.p, delta);
        B.axpy(nn, -1.0, P.p, Pold.p);
        delta = B.nrm2(nn, Pold.p) / 2.0;
        it++;
        if (o.trace) std::fprintf(stderr, "librint scf iter %d delta %.3e\n", it, delta);
    }
    out.iters = it;
    ctx.note_scf_purified(npurified);
    out.delta = delta;
    out.dE = dEn;
    out.P.resize(nn);
    hck(hipMemcpy(out.P.data(), P.p, sizeof(double) * nn, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    if (!converged()) {
        out.status = 1;
        char buf[160];
        std::snprintf(buf, sizeof buf, "SCF did not converge in %d cycles: |dP| = %.3e > %.3e", imax, delta, conv);
        out.msg = buf;
        return out;
    }
    // validity checks (reference src/scf.rs:572-600)
    const double tr = B.dot(nn, P.p, S.p);
    if (std::fabs(tr - nelec) > 1e-6) {
        out.status = 2;
        out.msg = "converged density has tr(PS) != nelec";
        return out;
    }
    B.gemm(n, false, false, 1.0, P.p, S.p, 0.0, T1.p);
    B.gemm(n, false, false, 1.0, T1.p, P.p, 0.0, T2.p);
    B.axpy(nn, -2.0, P.p, T2.p);
    if (B.amax(nn, T2.p) > 1e-6) {
        out.status = 2;
        out.msg = "converged density is not idempotent";
        return out;
    }
    ctx.fock_device(P.p, G.p, 0);
    out.E = Enuc + B.dot(nn, P.p, H.p) + 0.5 * B.dot(nn, P.p, G.p);
    out.status = 0;
    return out;
}

}  // namespace librint
