#pragma once

#include <string>
#include <vector>

#include "context.h"
#include "int1e.h"

namespace librint {

struct ScfOptions {
    int guess = 0;              // 0 identity (reference), 1 core Hamiltonian, 2 P0
    int diis = 0;               // Pulay DIIS on/off
    const double *P0 = nullptr; // starting density for guess == 2
    int trace = 0;
    double conv_e = 0.0;        // > 0: also require |E_k - E_(k-1)| <= conv_e
};

struct ScfOut {
    int status = -1;  // 0 ok, 1 not converged, 2 invalid density, 3 singular S
    int iters = 0;
    double E = 0.0;
    double delta = 0.0;
    double dE = 0.0;  // last energy change (conv_e mode)
    std::vector<double> P;
    std::string msg;
};

ScfOut scf_run(Context &ctx, const Int1eInput &in, int nelec, int imax, double conv, const ScfOptions &o);
// Q = P F P for symmetric host matrices (row- and column-major agree), on the GPU
void pfp_gpu(int n, const double *P, const double *F, double *Q);
// E[P] = tr(P H) + 1/2 tr(P G[P]) + E_nuc with G from the GPU Fock build
double energy_fast(Context &ctx, const Int1eInput &in, const double *P, const double *H);

}  // namespace librint
