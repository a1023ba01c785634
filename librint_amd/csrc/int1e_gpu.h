// One-electron integrals on the GPU (int1e_kernel.hip) and the host driver that
// stages a molecule for them.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "int1e.h"

namespace librint {

struct Shell1Dev {
    int l, nprim, pexp, pcoef, ao, pad;
    double r[3];
};

// (i, j) blocks of S, T and/or V (null pointers are skipped) for the ordered
// shell pairs `pairs` of class (li, lj); nuc[c] = {C, Z_C}.  Matrices nao x nao
// row-major, device pointers.
hipError_t launch_int1e(int li, int lj, const int2 *pairs, int npair, const Shell1Dev *sh, const double *env,
                        const double4 *nuc, int nnuc, int nao, double *S, double *T, double *V, hipStream_t st);

// Bra derivatives of the primitive integrals of the ordered pairs of class
// (li, lj) <= 2, weighted by W (W-weighted sums into gatom[3*natm] and
// denv[env index < env_l], any may be null) or, for typ 0 with dS non-null,
// as the matrix dS[(a*nao + b)*L + (e - s1)] for env entries s1 <= e < s1+L.
// shells' pad field holds the atom index.
hipError_t launch_int1e_deriv(int li, int lj, const int2 *pairs, int npair, const Shell1Dev *sh, const double *env,
                              const double4 *nuc, int nnuc, int nao, int typ, const double *W, double *gatom,
                              double *denv, long long env_l, double *dS, long long s1, long long L, hipStream_t st);

// Host driver: S, T, V (any may be null) into host buffers nao*nao, computed
// on the current GPU.  Throws on failure; l > 4 is rejected.
void int1e_gpu(const Int1eInput &in, double *S, double *T, double *V);
// Same into device buffers (nao*nao each, any may be null), on stream st.
void int1e_gpu_device(const Int1eInput &in, double *dS, double *dT, double *dV, hipStream_t st);

}  // namespace librint
