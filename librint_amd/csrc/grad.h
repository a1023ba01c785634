#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace librint {

// Adjoint two-electron kernel: one launch per angular class and side
// (0: bra centres a, b; 1: ket centres c, d) over the class's task list.
// Each task's fixed pair (JKTask::bra) is the differentiated pair: side 0
// takes the bra-major list of the Fock build (fixed canonical bra, window of
// canonical kets), side 1 a ket-major list of the same quartets (fixed
// canonical ket, window of up to 64 canonical bras of one bin).
//   gatom[3*natm]   += dE2/dR (nuclear coordinates)
//   dprim[2*slot]   += dE2/d(exponent), dprim[2*slot+1] += dE2/d(coefficient)
// where slot = shell_info[s].y + primitive index (one slot per primitive of
// each contracted shell), shell_info[s] = {atom, first slot}, coef[slot] the
// contraction coefficient as stored in env, pp_idx[pp] = primitive indices
// (on shell a, on shell b) of each primitive pair record.
hipError_t launch_grad(int la, int lb, int lc, int ld, int side, const JKTask *tasks, int ntask, const DevBasis &B,
                       const int2 *pp_idx, const int2 *shell_info, const double *coef, const double *D, int nao,
                       double *gatom, double *dprim, hipStream_t st);

} // namespace librint
