"""The reference's gradient API (reference python/librint/dscf.py:14-121) over
the GPU library: derivatives of the RHF energy with respect to the
basis-parameter slice env[s1:s2] (every primitive exponent and contraction
coefficient), plus nuclear-coordinate gradients, which the reference does
not provide.

The two-electron part (dRf) is the adjoint Rys kernel on the GPU; the
one-electron parts are host code.  ``mol`` is any object with pyscf-style
``_atm``, ``_bas``, ``_env`` and ``nelec``.
"""
import ctypes

import numpy as np

from ._bindings import library
from . import utils

# The adjoint kernels cover cartesian shells with l <= 4 (the reference's
# memory-safe domain, reference python/librint/dscf.py:11).
GRAD_LMAX = 4


def _require_grad_domain(mol):
    # same checks as reference python/librint/dscf.py:14-23, with this
    # library's angular-momentum limit
    lmax = max(int(b[1]) for b in np.asarray(mol._bas).reshape(-1, 8))
    omega = float(mol._env[8]) if len(mol._env) > 8 else 0.0
    if lmax > GRAD_LMAX or omega != 0.0:
        raise ValueError(
            f"librint_amd analytic 2e gradient supports only cartesian shells with l <= {GRAD_LMAX} "
            f"and no range separation (got max l = {lmax}, omega = {omega})")


def split(bas) -> tuple:
    """(s1, s2) of the differentiable env slice (reference src/utils.rs:152-175)."""
    mn, mx = -1, -1
    for b in np.asarray(bas).reshape(-1, 8):
        ngto, ex, cont = int(b[2]), int(b[5]), int(b[6])
        if mn == -1 or mn > ex:
            mn = ex
        if mx == -1:
            mx = cont + ngto
        elif mx < cont:
            mx = cont + ngto
    return mn, mx


def _call(fn, mol, P, shape):
    atm, bas, env, nelec = utils.prep(mol)
    P = np.ascontiguousarray(P, dtype=np.float64)
    out = fn(*utils.mol_args(atm, bas, env), P.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), P.size)
    if not out:
        raise RuntimeError("librint_amd gradient failed: " + utils.last_error())
    return utils.take(out, shape)


def grad(mol, P: np.ndarray) -> np.ndarray:
    """dHcore + dR (reference python/librint/dscf.py:26-37)."""
    _require_grad_domain(mol)
    s1, s2 = split(mol._bas)
    return _call(library.grad_c, mol, P, (s2 - s1,))


def dSu(mol) -> np.ndarray:
    """Uncontracted overlap derivative, shape (s2-s1, nao, nao) (reference python/librint/dscf.py:39-51)."""
    atm, bas, env, nelec = utils.prep(mol)
    n = utils.angl(bas, 0)
    s1, s2 = split(bas)
    out = library.dS_u(*utils.mol_args(atm, bas, env))
    if not out:
        raise RuntimeError("librint_amd dS_u failed: " + utils.last_error())
    return utils.take(out, (n, n, s2 - s1)).transpose(2, 0, 1)


def dSf(mol, P: np.ndarray) -> np.ndarray:
    """sum_ab (PFP)_ab dS_ab (reference python/librint/dscf.py:53-65)."""
    s1, s2 = split(mol._bas)
    return _call(library.dS_c, mol, P, (s2 - s1,))


def dHcoref(mol, P: np.ndarray) -> np.ndarray:
    """sum_ab P_ab d(T+V)_ab (reference python/librint/dscf.py:67-78)."""
    s1, s2 = split(mol._bas)
    return _call(library.dHcore_c, mol, P, (s2 - s1,))


def dRf(mol, P: np.ndarray) -> np.ndarray:
    """Two-electron part d/denv 1/2 tr(P G[P]) (reference python/librint/dscf.py:80-92)."""
    _require_grad_domain(mol)
    s1, s2 = split(mol._bas)
    return _call(library.dR_c, mol, P, (s2 - s1,))


def danalyticalf(mol, P: np.ndarray) -> np.ndarray:
    """dHcore + dR - dS/2 (reference python/librint/dscf.py:94-106)."""
    _require_grad_domain(mol)
    s1, s2 = split(mol._bas)
    return _call(library.danalytical_c, mol, P, (s2 - s1,))


def denergyf(mol, P: np.ndarray) -> np.ndarray:
    """Same as danalyticalf (reference python/librint/dscf.py:108-121)."""
    _require_grad_domain(mol)
    s1, s2 = split(mol._bas)
    return _call(library.denergy_c, mol, P, (s2 - s1,))


def grad_nuc(mol, P: np.ndarray) -> np.ndarray:
    """Nuclear-coordinate RHF gradient (natm, 3) at a converged density."""
    _require_grad_domain(mol)
    return _call(library.grad_nuc_c, mol, P, (len(mol._atm), 3))
