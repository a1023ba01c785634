#pragma once

#include <array>
#include <cstddef>
#include <vector>

namespace librint {

struct Shell1 {
    int l, nprim, pexp, pcoef, ao, atom;
    double r[3];
};

struct Int1eInput {
    std::vector<Shell1> shells;
    std::vector<double> env;
    std::vector<double> charges;  // |Z|: the nuclear attraction (reference src/g1e.rs:270-274)
    std::vector<double> zsigned;  // Z as given: the nuclear repulsion (reference src/scf.rs:946-953)
    std::vector<std::array<double, 3>> centers;
    int nao = 0;
};

Int1eInput make_int1e_input(const int *atm, size_t atm_l, const int *bas, size_t bas_l, const double *env, size_t env_l0);
double nuclear_repulsion(const Int1eInput &in);

// On the GPU (int1e_gpu.cpp): derivatives of E1 = sum_ab W_ab O_ab (W
// symmetric host matrix nao*nao; O = S (typ 0),
// T (1) or V (2)) with respect to every primitive exponent and contraction
// coefficient (added into denv[env index], entries below env_l) and to the
// nuclear coordinates (added into gatom[3*natm]; for V this includes the
// derivative with respect to the point-charge positions).  Either output may
// be null.
void int1e_grad(const Int1eInput &in, int typ, const double *W, double *gatom, double *denv, size_t env_l);
// Several such passes with one staging of the basis on the GPU.
struct Int1eGradPass {
    int typ;
    const double *W;
    double *gatom, *denv;
};
void int1e_grad_passes(const Int1eInput &in, const std::vector<Int1eGradPass> &passes, size_t env_l);
// d/dO_ab of every overlap element: dS[(a*nao + b)*L + (e - s1)] for env
// entries s1 <= e < s1 + L.
void overlap_param_derivs(const Int1eInput &in, size_t s1, size_t L, double *dS);
// gradient of the point-charge repulsion (added into gatom[3*natm])
void nuclear_repulsion_grad(const Int1eInput &in, double *gatom);

}  // namespace librint
