// Kernels of the learned SCF guess: a small MLP (BF16 weights and
// activations, FP32 accumulation and master weights) trained and evaluated
// on the GPU.
//
// gemm_bf16: C = act(op(A) op(B) + bias) on the matrix cores
// (v_mfma_f32_16x16x16_bf16), op = identity or transpose, so that one kernel
// serves the forward pass (X W), the input gradient (dZ W^T) and the weight
// gradient (X^T dZ).  A workgroup of 4 wavefronts owns a 64 x 64 tile of C;
// each wavefront a 32 x 32 quarter (2 x 2 MFMA tiles).  K advances in steps
// of 16 with both operand tiles staged in LDS k-contiguous, so that every
// lane fetches its 4 consecutive k values of a row (A) or column (B) with one
// 8-byte LDS read.  Out-of-range rows, columns and k are zero-filled.
//
// MFMA 16x16x16 operand layout (lane l = 0..63): A(row l%16, k 4(l/16)+i),
// B(k 4(l/16)+i, column l%16), result D(row 4(l/16)+i, column l%16), i = 0..3.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.h"

namespace librint {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even (inputs are finite)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

constexpr int GT = 64;   // tile edge
constexpr int GK = 16;   // k step
constexpr int GKP = 20;  // padded k stride in LDS (40 bytes: 8-byte aligned rows)

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, const uint16_t *__restrict__ A, int lda,
                                                        const uint16_t *__restrict__ B, int ldb, float *__restrict__ C,
                                                        int ldc, const float *__restrict__ bias, int act, int accumulate,
                                                        uint16_t *__restrict__ Cb, int ldcb)
{
    __shared__ __align__(16) uint16_t As[GT * GKP];  // [m][k]
    __shared__ __align__(16) uint16_t Bs[GT * GKP];  // [n][k]
    const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < K; k0 += GK) {
        // stage 64 x 16 of A and of B: 4 elements per thread each
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int idx = tid + 256 * e;  // 0..1023
            int r, k;
            if (TA) {  // consecutive threads along m (contiguous in memory)
                r = idx % GT;
                k = idx / GT;
            } else {   // consecutive threads along k
                r = idx / GK;
                k = idx % GK;
            }
            const int gm = m0 + r, gk = k0 + k;
            uint16_t v = 0;
            if (gm < M && gk < K) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[r * GKP + k] = v;
            int c, kb;
            if (TB) {  // B(k, n) = B[n*ldb + k]: contiguous along k
                c = idx / GK;
                kb = idx % GK;
            } else {   // B(k, n) = B[k*ldb + n]: contiguous along n
                c = idx % GT;
                kb = idx / GT;
            }
            const int gn = n0 + c, gk2 = k0 + kb;
            uint16_t w = 0;
            if (gn < N && gk2 < K) w = TB ? B[(size_t)gn * ldb + gk2] : B[(size_t)gk2 * ldb + gn];
            Bs[c * GKP + kb] = w;
        }
        __syncthreads();
        const int kq = 4 * (lane >> 4), rl = lane & 15;
        v4s a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            a[i] = *reinterpret_cast<const v4s *>(&As[(wm + 16 * i + rl) * GKP + kq]);
            b[i] = *reinterpret_cast<const v4s *>(&Bs[(wn + 16 * i + rl) * GKP + kq]);
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // epilogue: lane holds rows 4(l/16)+q, column l%16 of each 16 x 16 tile
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int gm = m0 + wm + 16 * i + 4 * (lane >> 4) + q;
                const int gn = n0 + wn + 16 * j + (lane & 15);
                if (gm >= M || gn >= N) continue;
                float v = acc[i][j][q];
                if (accumulate) v += C[(size_t)gm * ldc + gn];
                if (bias) v += bias[gn];
                if (act == 1) v = v > 0.f ? v : 0.f;
                if (C) C[(size_t)gm * ldc + gn] = v;
                if (Cb) Cb[(size_t)gm * ldcb + gn] = f2bf(v);
            }
}

hipError_t launch_gemm_bf16(int M, int N, int K, const uint16_t *A, int lda, int transA, const uint16_t *B, int ldb,
                            int transB, float *C, int ldc, const float *bias, int act, int accumulate, uint16_t *Cb,
                            int ldcb, hipStream_t st)
{
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    if (!C && !Cb) return hipErrorInvalidValue;
    if (accumulate && !C) return hipErrorInvalidValue;
    const dim3 grid((M + GT - 1) / GT, (N + GT - 1) / GT);
    if (grid.y > 65535) return hipErrorInvalidValue;
#define L(ta, tb)                                                                                                      \
    hipLaunchKernelGGL((gemm_bf16_kernel<ta, tb>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias, act, \
                       accumulate, Cb, ldcb)
    if (transA && transB) L(true, true);
    else if (transA) L(true, false);
    else if (transB) L(false, true);
    else L(false, false);
#undef L
    return hipGetLastError();
}

// dpre = dz * (h > 0) (ReLU backward through the stored activation h), as
// FP32 and BF16
__global__ void relu_bwd_kernel(size_t n, const float *__restrict__ dz, const float *__restrict__ h,
                                float *__restrict__ dpre, uint16_t *__restrict__ dpre_b)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = h[i] > 0.f ? dz[i] : 0.f;
    if (dpre) dpre[i] = v;
    if (dpre_b) dpre_b[i] = f2bf(v);
}

hipError_t launch_relu_bwd(size_t n, const float *dz, const float *h, float *dpre, uint16_t *dpre_b, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(relu_bwd_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, dz, h, dpre, dpre_b);
    return hipGetLastError();
}

// out[c] (+)= sum_r x[r*ld + c]: one workgroup per 64 columns, 4 row stripes
__global__ __launch_bounds__(256) void colsum_kernel(int rows, int cols, const float *__restrict__ x, int ld,
                                                     float *__restrict__ out, int accumulate)
{
    __shared__ float part[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), w = threadIdx.x >> 6;
    float s = 0.f;
    if (c < cols)
        for (int r = w; r < rows; r += 4) s += x[(size_t)r * ld + c];
    part[w][threadIdx.x & 63] = s;
    __syncthreads();
    if (w == 0 && c < cols) {
        const float v = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
        out[c] = accumulate ? out[c] + v : v;
    }
}

hipError_t launch_colsum(int rows, int cols, const float *x, int ld, float *out, int accumulate, hipStream_t st)
{
    if (cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(colsum_kernel, dim3((cols + 63) / 64), dim3(256), 0, st, rows, cols, x, ld, out, accumulate);
    return hipGetLastError();
}

// masked mean squared error over the first `cols` columns of y (row stride
// ldy) against t (row stride ldt): dy = 2 mask (y - t) / scale (BF16 and FP32;
// columns >= cols of dy are zeroed), loss[0] += sum mask (y - t)^2 / scale
__global__ __launch_bounds__(256) void mse_kernel(int rows, int cols, const float *__restrict__ y, int ldy,
                                                  const float *__restrict__ t, const float *__restrict__ mask, int ldt,
                                                  float scale, float *__restrict__ dy, uint16_t *__restrict__ dy_b,
                                                  double *__restrict__ loss)
{
    __shared__ double red[256];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)rows * ldy;
    double l = 0.0;
    if (i < n) {
        const int r = (int)(i / ldy), c = (int)(i % ldy);
        float g = 0.f;
        if (c < cols) {
            const float m = mask ? mask[(size_t)r * ldt + c] : 1.f;
            const float d = y[i] - t[(size_t)r * ldt + c];
            g = 2.f * m * d / scale;
            l = (double)m * d * d / scale;
        }
        if (dy) dy[i] = g;
        if (dy_b) dy_b[i] = f2bf(g);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) atomicAdd(loss, red[0]);
}

hipError_t launch_mse(int rows, int cols, const float *y, int ldy, const float *t, const float *mask, int ldt,
                      float scale, float *dy, uint16_t *dy_b, double *loss, hipStream_t st)
{
    const size_t n = (size_t)rows * ldy;
    if (n == 0) return hipSuccess;
    if (cols > ldy || cols > ldt) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mse_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rows, cols, y, ldy, t, mask,
                       ldt, scale, dy, dy_b, loss);
    return hipGetLastError();
}

// Adam step on FP32 master weights, refreshing the BF16 copy
__global__ void adam_kernel(size_t n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
                            uint16_t *__restrict__ wb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

hipError_t launch_adam(size_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                       int step, uint16_t *wb, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    if (step < 1) return hipErrorInvalidValue;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, w, g, m, v, lr, b1, b2, eps,
                       c1, c2, wb);
    return hipGetLastError();
}

__global__ void f32_to_bf16_kernel(size_t n, const float *__restrict__ x, uint16_t *__restrict__ y)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

hipError_t launch_f32_to_bf16(size_t n, const float *x, uint16_t *y, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, x, y);
    return hipGetLastError();
}

} // namespace librint
