// Two-electron kernels for angular classes with an f or g shell (l = 3, 4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "kernels.h"

namespace librint {

// Shape of one angular class, known at run time: (la <= lb | lc <= ld) as the
// pair records store them, g index i + k*dk + l*dl + j*dj over the raised
// momenta, nroots Rys roots handled rb at a time (the LDS holds 3*rb chains).
struct HLShape {
    int la, lb, lc, ld;
    int nfa, nfb, nfc, nfd, nf;
    int nmax, mmax, nroots;
    int dk, dl, dj, gs;
    int rb;
};

constexpr int HL_LMAX = 4;

HLShape hl_shape(int la, int lb, int lc, int ld);
// true when the class has a shell with l > 2 (served by the kernels here)
inline bool hl_class(int la, int lb, int lc, int ld) { return la > 2 || lb > 2 || lc > 2 || ld > 2; }
// doubles of global scratch one workgroup uses (one contracted block)
inline size_t hl_scratch_per_group(const HLShape &S) { return (size_t)S.nf; }

// Fock build over JKTasks (bra pair x up to 64 kets of one bin); M += the
// A - K/2 contributions of every task as the other Fock kernels do.  scr
// holds scr_len doubles; the launch uses min(ntask, scr_len / nf, 8192)
// workgroups, each looping over tasks.
hipError_t launch_hl_jk(const HLShape &S, const JKTask *tasks, int ntask, const unsigned char *flags, int accept,
                        const DevBasis &B, const double *D, double *M, int nao, double *scr, size_t scr_len,
                        hipStream_t st);
// Quartet lists, modes as launch_list: 0 block to out[t.out ...], 1 Schwarz
// sqrt(max |(ab|ab)|) to out[t.out], 2 scatter into the nao^4 tensor.
hipError_t launch_hl_list(const HLShape &S, const ListTask *tasks, int ntask, const DevBasis &B, double *out, int mode,
                          int nao, double *scr, size_t scr_len, hipStream_t st);

// Shape of the adjoint's recurrence for the class (xa xb | ya yb) with the
// pair (xa, xb) differentiated (momentum raised by 2 on both its centres).
HLShape hl_grad_shape(int xa, int xb, int ya, int yb);
// Adjoint of E2 = 1/2 tr(P G[P]) over JKTasks of a class with an f or g
// shell (side 0: the task's bra pair is differentiated, side 1: its fixed
// pair is the canonical ket of a ket-major list); outputs and per-primitive
// slots as launch_grad (grad.h).  scr as launch_hl_jk (one weight block of
// nf doubles per workgroup).
hipError_t launch_hl_grad(int la, int lb, int lc, int ld, int side, const JKTask *tasks, int ntask, const DevBasis &B,
                          const int2 *pp_idx, const int2 *shell_info, const double *coef, const double *D, int nao,
                          double *gatom, double *dprim, double *scr, size_t scr_len, hipStream_t st);

}  // namespace librint
