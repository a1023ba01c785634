// Cartesian -> real spherical transform for l <= 2 (reference
// src/cart2sph.rs:88380-88530; the l = 1 transform is the identity in the
// px, py, pz order, as there).  The d rows are the real solid harmonics of
// degree 2 in the order m = -2..2 (xy, yz, z^2, xz, x^2 - y^2) over the
// cartesian order xx, xy, xz, yy, yz, zz, with the normalisation of
// Y_2m * r^2:
//   xy, yz, xz:  sqrt(15 / (4 pi))
//   z^2:         sqrt(5 / (16 pi)) * (2 zz - xx - yy)
//   x^2 - y^2:   sqrt(15 / (16 pi)) * (xx - yy)
#include "cart2sph.h"

#include <cmath>
#include <stdexcept>

namespace librint {

std::vector<C2SRow> c2s_rows(const int *bas, size_t bas_l, int *ncart)
{
    const double a = std::sqrt(15.0 / (4.0 * M_PI));
    const double b = std::sqrt(5.0 / (16.0 * M_PI));
    const double c = std::sqrt(15.0 / (16.0 * M_PI));
    std::vector<C2SRow> rows;
    int cart = 0;
    for (size_t s = 0; s + 8 <= bas_l; s += 8) {
        const int l = bas[s + 1], nctr = bas[s + 3];
        if (l < 0 || l > 2) throw std::invalid_argument("spherical output covers l <= 2");
        for (int k = 0; k < nctr; k++) {
            if (l == 0) {
                rows.push_back(C2SRow{1, {cart, 0, 0}, {1.0, 0.0, 0.0}});
                cart += 1;
            } else if (l == 1) {
                for (int m = 0; m < 3; m++) rows.push_back(C2SRow{1, {cart + m, 0, 0}, {1.0, 0.0, 0.0}});
                cart += 3;
            } else {
                // cartesian order xx 0, xy 1, xz 2, yy 3, yz 4, zz 5
                rows.push_back(C2SRow{1, {cart + 1, 0, 0}, {a, 0.0, 0.0}});
                rows.push_back(C2SRow{1, {cart + 4, 0, 0}, {a, 0.0, 0.0}});
                rows.push_back(C2SRow{3, {cart + 0, cart + 3, cart + 5}, {-b, -b, 2.0 * b}});
                rows.push_back(C2SRow{1, {cart + 2, 0, 0}, {a, 0.0, 0.0}});
                rows.push_back(C2SRow{2, {cart + 0, cart + 3, 0}, {c, -c, 0.0}});
                cart += 6;
            }
        }
    }
    if (ncart) *ncart = cart;
    return rows;
}

__global__ void c2s2_kernel(const C2SRow *__restrict__ rows, int nsph, const double *__restrict__ in, int ncart,
                            double *__restrict__ out)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ns = (size_t)nsph;
    if (t >= ns * ns) return;
    const C2SRow r0 = rows[t / ns], r1 = rows[t % ns];
    const size_t nc = (size_t)ncart;
    double acc = 0.0;
    for (int i = 0; i < r0.n; i++)
        for (int j = 0; j < r1.n; j++) acc += r0.c[i] * r1.c[j] * in[(size_t)r0.cart[i] * nc + r1.cart[j]];
    out[t] = acc;
}

__global__ void c2s4_kernel(const C2SRow *__restrict__ rows, int nsph, const double *__restrict__ in, int ncart,
                            double *__restrict__ out)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ns = (size_t)nsph;
    if (t >= ns * ns * ns * ns) return;
    const C2SRow r0 = rows[t / (ns * ns * ns)], r1 = rows[(t / (ns * ns)) % ns], r2 = rows[(t / ns) % ns],
                 r3 = rows[t % ns];
    const size_t nc = (size_t)ncart;
    double acc = 0.0;
    for (int i = 0; i < r0.n; i++)
        for (int j = 0; j < r1.n; j++)
            for (int k = 0; k < r2.n; k++)
                for (int l = 0; l < r3.n; l++)
                    acc += r0.c[i] * r1.c[j] * r2.c[k] * r3.c[l] *
                           in[(((size_t)r0.cart[i] * nc + r1.cart[j]) * nc + r2.cart[k]) * nc + r3.cart[l]];
    out[t] = acc;
}

hipError_t launch_c2s(int ndim, const C2SRow *rows, int nsph, const double *in, int ncart, double *out,
                      hipStream_t st)
{
    if (nsph <= 0) return hipSuccess;
    const size_t ns = (size_t)nsph;
    if (ndim == 2) {
        const size_t m = ns * ns;
        hipLaunchKernelGGL(c2s2_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, rows, nsph, in, ncart, out);
    } else if (ndim == 4) {
        const size_t m = ns * ns * ns * ns;
        if ((m + 255) / 256 > 0x7fffffffull) return hipErrorInvalidValue;
        hipLaunchKernelGGL(c2s4_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, rows, nsph, in, ncart, out);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace librint
