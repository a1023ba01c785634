// One-electron input staging (shells, env, point charges) and the
// point-charge repulsion energy and gradient.  The one-electron integrals and
// their derivatives are computed on the GPU (int1e_kernel.hip,
// int1e_gpu.cpp).
#include <algorithm>
#include <cmath>
#include <vector>

#include "int1e.h"

namespace librint {

Int1eInput make_int1e_input(const int *atm, size_t atm_l, const int *bas, size_t bas_l, const double *env, size_t env_l0)
{
    Int1eInput in;
    in.env.assign(env, env + 0);
    size_t env_need = 0;
    const int nb = (int)(bas_l / 8);
    int ao = 0;
    for (int s = 0; s < nb; s++) {
        const int *b = bas + 8 * s;
        for (int c = 0; c < b[3]; c++) {
            Shell1 sh;
            sh.l = b[1];
            sh.nprim = b[2];
            sh.pexp = b[5];
            sh.pcoef = b[6] + c * b[2];
            sh.atom = b[0];
            const int pr = atm[6 * b[0] + 1];
            for (int d = 0; d < 3; d++) sh.r[d] = env[pr + d];
            sh.ao = ao;
            ao += (sh.l + 1) * (sh.l + 2) / 2;
            env_need = std::max(env_need, (size_t)std::max(sh.pexp, sh.pcoef) + sh.nprim);
            in.shells.push_back(sh);
        }
    }
    for (size_t a = 0; a < atm_l / 6; a++) {
        const int pr = atm[6 * a + 1];
        env_need = std::max(env_need, (size_t)pr + 3);
        // point nuclei only: the Gaussian charge model (nuc_mod 2 with a
        // zeta > 0, reference src/g1e.rs:226-239) and the fractional-charge
        // model (nuc_mod 3) are refused rather than treated as point charges
        const int mod = atm[6 * a + 2];
        if (mod == 3) throw std::invalid_argument("fractional nuclear charges (nuc_mod 3) are not supported");
        if (mod == 2) {
            const int pz = atm[6 * a + 3];
            if (pz >= 0 && (size_t)pz < env_l0 && env[pz] > 0.0)
                throw std::invalid_argument("Gaussian nuclear charge distributions (nuc_mod 2, zeta > 0) are not supported");
        }
        in.charges.push_back(std::fabs((double)atm[6 * a]));
        in.zsigned.push_back((double)atm[6 * a]);
        in.centers.push_back({env[pr], env[pr + 1], env[pr + 2]});
    }
    in.env.assign(env, env + env_need);
    in.nao = ao;
    return in;
}

void nuclear_repulsion_grad(const Int1eInput &in, double *gatom)
{
    const size_t na = in.charges.size();
    for (size_t i = 0; i < na; i++)
        for (size_t j = 0; j < na; j++) {
            if (i == j) continue;
            double d[3], r2 = 0;
            for (int k = 0; k < 3; k++) {
                d[k] = in.centers[i][k] - in.centers[j][k];
                r2 += d[k] * d[k];
            }
            const double f = in.zsigned[i] * in.zsigned[j] / (r2 * std::sqrt(r2));
            for (int k = 0; k < 3; k++) gatom[3 * i + k] -= f * d[k];
        }
}

double nuclear_repulsion(const Int1eInput &in)
{ // reference src/scf.rs:946-953 (point charges, signed Z as given)
    double E = 0;
    for (size_t i = 0; i < in.charges.size(); i++)
        for (size_t j = 0; j < i; j++) {
            double r2 = 0;
            for (int d = 0; d < 3; d++) r2 += (in.centers[i][d] - in.centers[j][d]) * (in.centers[i][d] - in.centers[j][d]);
            E += in.zsigned[i] * in.zsigned[j] / std::sqrt(r2);
        }
    return E;
}

}  // namespace librint
