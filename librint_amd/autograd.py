"""RHF energy as a differentiable torch function of the nuclear coordinates
and the Gaussian exponents.

Forward: converged RHF on the GPU (persistent context, DIIS) for the given
coordinates (bohr) and exponents.  Backward: the analytic gradient at the
converged density -- the adjoint Rys kernel for the two-electron part, host
one-electron derivatives, the Pulay term -1/2 tr(PFP dS) and the nuclear
repulsion -- through ``librint_ctx_gradient``.  Exponent derivatives are at
fixed env coefficients, the convention of the reference's basis-parameter
gradients (reference src/dscf.rs:526-537).
"""
import ctypes

import numpy as np
import torch

from ._bindings import library
from . import utils
from .context import FockContext
from .mol import Mole


def exponent_slots(mol):
    """Sorted env indices of every primitive exponent of the basis."""
    idx = set()
    for b in np.asarray(mol._bas).reshape(-1, 8):
        idx.update(range(int(b[5]), int(b[5]) + int(b[2])))
    return np.array(sorted(idx), dtype=np.int64)


def with_params(mol, coords, exps, slots):
    """Copy of mol with atom coordinates (natm, 3) and env[slots] = exps."""
    env = np.array(mol._env, dtype=np.float64, copy=True)
    for a, row in enumerate(np.asarray(mol._atm).reshape(-1, 6)):
        env[int(row[1]):int(row[1]) + 3] = coords[a]
    env[slots] = exps
    return Mole(np.array(mol._atm, copy=True), np.array(mol._bas, copy=True), env, mol.nelectron, mol.name)


def rhf(mol, conv=1e-10, imax=200, schwarz_tol=1e-14, P0=None):
    """Converged RHF on the GPU: (energy, density, iterations, context)."""
    ctx = FockContext(mol, schwarz_tol)
    if P0 is None:
        E, P, it = ctx.scf(imax=imax, conv=conv, guess=1, diis=True)
    else:
        E, P, it = ctx.scf(imax=imax, conv=conv, guess=2, diis=True, P0=P0)
    if not np.isfinite(E):
        raise RuntimeError("RHF did not converge: " + utils.last_error())
    return E, P, it, ctx


def gradient(ctx, P, parts=False):
    """(dE/dR (natm, 3), dE/denv (env_l,)) at the converged density P."""
    P = np.ascontiguousarray(P, dtype=np.float64)
    natm = len(ctx._keep[0]) // 6 if ctx._keep[0].ndim == 1 else ctx._keep[0].shape[0]
    env_l = ctx._keep[2].size
    gn = np.zeros(3 * natm)
    de = np.zeros(env_l)
    dp = ctypes.POINTER(ctypes.c_double)
    pb = np.zeros(4 * 3 * natm + 3 * env_l) if parts else None
    rc = library.librint_ctx_gradient(ctx.h, P.ctypes.data_as(dp), gn.ctypes.data_as(dp), de.ctypes.data_as(dp),
                                      pb.ctypes.data_as(dp) if parts else None)
    if rc != 0:
        raise RuntimeError("librint gradient failed: " + utils.last_error())
    if not parts:
        return gn.reshape(natm, 3), de
    m = 3 * natm
    names = ("gH", "gR", "gS", "gN")
    out = {k: pb[i * m:(i + 1) * m].reshape(natm, 3) for i, k in enumerate(names)}
    for i, k in enumerate(("dH", "dR", "dS")):
        out[k] = pb[4 * m + i * env_l:4 * m + (i + 1) * env_l]
    return gn.reshape(natm, 3), de, out


class RHFEnergy(torch.autograd.Function):
    """E = RHFEnergy.apply(coords (natm, 3) bohr, exps (nexp,), mol, slots)."""

    @staticmethod
    def forward(ctx, coords, exps, mol, slots, conv=1e-10):
        m = with_params(mol, coords.detach().cpu().double().numpy(), exps.detach().cpu().double().numpy(), slots)
        E, P, it, fctx = rhf(m, conv=conv)
        gn, de = gradient(fctx, P)
        fctx.close()
        ctx.save_for_backward(torch.as_tensor(gn, dtype=coords.dtype, device=coords.device),
                              torch.as_tensor(de[slots], dtype=exps.dtype, device=exps.device))
        ctx.iterations = it
        return torch.tensor(E, dtype=coords.dtype, device=coords.device)

    @staticmethod
    def backward(ctx, gE):
        gn, ge = ctx.saved_tensors
        return gE * gn, gE * ge, None, None, None
