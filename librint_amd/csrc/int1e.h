#pragma once

#include <array>
#include <cstddef>
#include <vector>

namespace librint {

struct Shell1 {
    int l, nprim, pexp, pcoef, ao;
    double r[3];
};

struct Int1eInput {
    std::vector<Shell1> shells;
    std::vector<double> env;
    std::vector<double> charges;
    std::vector<std::array<double, 3>> centers;
    int nao = 0;
};

Int1eInput make_int1e_input(const int *atm, size_t atm_l, const int *bas, size_t bas_l, const double *env);
// typ 0 overlap, 1 kinetic, 2 nuclear attraction; R[mu*nao + nu]
void int1e_matrix(const Int1eInput &in, int typ, double *R);
double nuclear_repulsion(const Int1eInput &in);

}  // namespace librint
