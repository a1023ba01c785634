// Cartesian -> real spherical transform of integral matrices and tensors
// (l <= 4), the transform of reference src/cart2sph.rs:88380-88530.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

namespace librint {

// One spherical function: up to 6 cartesian components with coefficients.
constexpr int C2S_MAXTERM = 6;
struct C2SRow {
    int n;
    int cart[C2S_MAXTERM];
    double c[C2S_MAXTERM];
};

// Rows for every spherical function of a basis (bas in libcint layout: l and
// nctr per shell), in spherical AO order; ncart receives the cartesian count.
// Throws for l > 4.
std::vector<C2SRow> c2s_rows(const int *bas, size_t bas_l, int *ncart);

// out[sph^ndim] = sum over cartesian components of prod_k c_k * in[cart^ndim]
// (ndim 2 or 4; row-major, first index slowest), device pointers.
hipError_t launch_c2s(int ndim, const C2SRow *rows, int nsph, const double *in, int ncart, double *out,
                      hipStream_t st);

}  // namespace librint
