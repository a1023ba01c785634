// Cartesian -> real spherical transform for l <= 4 (reference
// src/cart2sph.rs:88380-88530; the l = 1 transform is the identity in the
// px, py, pz order, as there).  Each spherical function of degree l >= 2 is
// the real solid harmonic r^l Y_lm (Y_lm normalised on the sphere, no
// Condon-Shortley phase), m = -l..l, written over the cartesian monomials
// of the shell in libcint order (x power descending, then y):
//   d:  m=-2 sqrt(15/4pi) xy, m=-1 sqrt(15/4pi) yz,
//       m=0 sqrt(5/16pi) (2z^2 - x^2 - y^2), m=1 sqrt(15/4pi) xz,
//       m=2 sqrt(15/16pi) (x^2 - y^2)
//   f:  m=-3 sqrt(35/32pi) (3x^2 - y^2) y, m=-2 sqrt(105/4pi) xyz,
//       m=-1 sqrt(21/32pi) y (4z^2 - x^2 - y^2),
//       m=0 sqrt(7/16pi) z (2z^2 - 3x^2 - 3y^2),
//       m=1 sqrt(21/32pi) x (4z^2 - x^2 - y^2),
//       m=2 sqrt(105/16pi) z (x^2 - y^2), m=3 sqrt(35/32pi) (x^2 - 3y^2) x
//   g:  m=-4 3/4 sqrt(35/pi) xy (x^2 - y^2),
//       m=-3 3/4 sqrt(35/2pi) (3x^2 - y^2) yz,
//       m=-2 3/4 sqrt(5/pi) xy (6z^2 - x^2 - y^2),
//       m=-1 3/4 sqrt(5/2pi) yz (4z^2 - 3x^2 - 3y^2),
//       m=0 3/16 sqrt(1/pi) (8z^4 + 3x^4 + 3y^4 + 6x^2y^2 - 24x^2z^2 - 24y^2z^2),
//       m=1 3/4 sqrt(5/2pi) xz (4z^2 - 3x^2 - 3y^2),
//       m=2 3/8 sqrt(5/pi) (x^2 - y^2)(6z^2 - x^2 - y^2),
//       m=3 3/4 sqrt(35/2pi) (x^2 - 3y^2) xz,
//       m=4 3/16 sqrt(35/pi) (x^4 - 6x^2y^2 + y^4)
#include "cart2sph.h"

#include <cmath>
#include <stdexcept>

namespace librint {

namespace {

// position of x^i y^j z^k (i + j + k = l) in the libcint cartesian order
int cart_pos(int l, int i, int j)
{
    const int row = l - i;  // rows of decreasing x power hold row + 1 entries
    return row * (row + 1) / 2 + (row - j);
}

struct Mono {
    int i, j, k;
    double c;
};

C2SRow make_row(int l, int base, std::initializer_list<Mono> terms, double scale)
{
    C2SRow r{};
    for (const Mono &t : terms) {
        if (t.i + t.j + t.k != l || r.n >= C2S_MAXTERM) throw std::logic_error("bad spherical row");
        r.cart[r.n] = base + cart_pos(l, t.i, t.j);
        r.c[r.n] = scale * t.c;
        r.n++;
    }
    return r;
}

void push_rows(std::vector<C2SRow> &rows, int l, int base)
{
    const double pi = M_PI;
    if (l == 0) {
        rows.push_back(make_row(0, base, {{0, 0, 0, 1.0}}, 1.0));
    } else if (l == 1) {
        rows.push_back(make_row(1, base, {{1, 0, 0, 1.0}}, 1.0));
        rows.push_back(make_row(1, base, {{0, 1, 0, 1.0}}, 1.0));
        rows.push_back(make_row(1, base, {{0, 0, 1, 1.0}}, 1.0));
    } else if (l == 2) {
        const double a = std::sqrt(15.0 / (4.0 * pi)), b = std::sqrt(5.0 / (16.0 * pi)),
                     c = std::sqrt(15.0 / (16.0 * pi));
        rows.push_back(make_row(2, base, {{1, 1, 0, 1.0}}, a));
        rows.push_back(make_row(2, base, {{0, 1, 1, 1.0}}, a));
        rows.push_back(make_row(2, base, {{2, 0, 0, -1.0}, {0, 2, 0, -1.0}, {0, 0, 2, 2.0}}, b));
        rows.push_back(make_row(2, base, {{1, 0, 1, 1.0}}, a));
        rows.push_back(make_row(2, base, {{2, 0, 0, 1.0}, {0, 2, 0, -1.0}}, c));
    } else if (l == 3) {
        const double c3 = std::sqrt(35.0 / (32.0 * pi)), c2a = std::sqrt(105.0 / (4.0 * pi)),
                     c2b = std::sqrt(105.0 / (16.0 * pi)), c1 = std::sqrt(21.0 / (32.0 * pi)),
                     c0 = std::sqrt(7.0 / (16.0 * pi));
        rows.push_back(make_row(3, base, {{2, 1, 0, 3.0}, {0, 3, 0, -1.0}}, c3));
        rows.push_back(make_row(3, base, {{1, 1, 1, 1.0}}, c2a));
        rows.push_back(make_row(3, base, {{0, 1, 2, 4.0}, {2, 1, 0, -1.0}, {0, 3, 0, -1.0}}, c1));
        rows.push_back(make_row(3, base, {{0, 0, 3, 2.0}, {2, 0, 1, -3.0}, {0, 2, 1, -3.0}}, c0));
        rows.push_back(make_row(3, base, {{1, 0, 2, 4.0}, {3, 0, 0, -1.0}, {1, 2, 0, -1.0}}, c1));
        rows.push_back(make_row(3, base, {{2, 0, 1, 1.0}, {0, 2, 1, -1.0}}, c2b));
        rows.push_back(make_row(3, base, {{3, 0, 0, 1.0}, {1, 2, 0, -3.0}}, c3));
    } else if (l == 4) {
        const double c4a = 0.75 * std::sqrt(35.0 / pi), c4b = 3.0 / 16.0 * std::sqrt(35.0 / pi),
                     c3 = 0.75 * std::sqrt(35.0 / (2.0 * pi)), c2a = 0.75 * std::sqrt(5.0 / pi),
                     c2b = 0.375 * std::sqrt(5.0 / pi), c1 = 0.75 * std::sqrt(5.0 / (2.0 * pi)),
                     c0 = 3.0 / 16.0 * std::sqrt(1.0 / pi);
        rows.push_back(make_row(4, base, {{3, 1, 0, 1.0}, {1, 3, 0, -1.0}}, c4a));
        rows.push_back(make_row(4, base, {{2, 1, 1, 3.0}, {0, 3, 1, -1.0}}, c3));
        rows.push_back(make_row(4, base, {{1, 1, 2, 6.0}, {3, 1, 0, -1.0}, {1, 3, 0, -1.0}}, c2a));
        rows.push_back(make_row(4, base, {{0, 1, 3, 4.0}, {2, 1, 1, -3.0}, {0, 3, 1, -3.0}}, c1));
        rows.push_back(make_row(4, base,
                                {{0, 0, 4, 8.0}, {4, 0, 0, 3.0}, {0, 4, 0, 3.0}, {2, 2, 0, 6.0}, {2, 0, 2, -24.0},
                                 {0, 2, 2, -24.0}},
                                c0));
        rows.push_back(make_row(4, base, {{1, 0, 3, 4.0}, {3, 0, 1, -3.0}, {1, 2, 1, -3.0}}, c1));
        rows.push_back(make_row(4, base, {{2, 0, 2, 6.0}, {0, 2, 2, -6.0}, {4, 0, 0, -1.0}, {0, 4, 0, 1.0}}, c2b));
        rows.push_back(make_row(4, base, {{3, 0, 1, 1.0}, {1, 2, 1, -3.0}}, c3));
        rows.push_back(make_row(4, base, {{4, 0, 0, 1.0}, {2, 2, 0, -6.0}, {0, 4, 0, 1.0}}, c4b));
    } else {
        throw std::invalid_argument("spherical output covers l <= 4");
    }
}

}  // namespace

std::vector<C2SRow> c2s_rows(const int *bas, size_t bas_l, int *ncart)
{
    std::vector<C2SRow> rows;
    int cart = 0;
    for (size_t s = 0; s + 8 <= bas_l; s += 8) {
        const int l = bas[s + 1], nctr = bas[s + 3];
        if (l < 0 || l > 4) throw std::invalid_argument("spherical output covers l <= 4");
        for (int k = 0; k < nctr; k++) {
            push_rows(rows, l, cart);
            cart += (l + 1) * (l + 2) / 2;
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
