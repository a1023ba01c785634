"""The reference's Python API (reference python/librint/scf.py:7-98) over the
GPU library: int1e, int2e, density, energy, scf with the same arguments,
return shapes and error behaviour.  ``mol`` is any object with pyscf-style
``_atm``, ``_bas``, ``_env`` and ``nelec`` (e.g. librint_amd.mol.Mole)."""
import ctypes

import numpy as np

from ._bindings import library
from . import utils


def int1e(mol, typei: str = 'ovlp', coord: str = 'cart') -> np.ndarray:
    atm, bas, env, nelec = utils.prep(mol)
    flag = {'ovlp': 0, 'kin': 1, 'nuc': 2}.get(typei)
    if flag is None:
        print("integral type does not exist: ovlp, kin, nuc")
        return None
    c = {'cart': 0, 'sph': 1}.get(coord)
    if c is None:
        print("coordinate type does not exist: cart, sph")
        return None
    n = utils.angl(bas, c)
    R = library.int1e_c(*utils.mol_args(atm, bas, env), c, flag)
    if not R:
        raise RuntimeError("librint int1e failed: " + utils.last_error())
    return utils.take(R, (n, n))


def int2e(mol, coord: str = 'cart') -> np.ndarray:
    atm, bas, env, nelec = utils.prep(mol)
    c = {'cart': 0, 'sph': 1}.get(coord)
    if c is None:
        print("coordinate type does not exist: cart, sph")
        return None
    n = utils.angl(bas, c)
    R = library.int2e_c(*utils.mol_args(atm, bas, env), c)
    if not R:
        raise RuntimeError("librint int2e failed: " + utils.last_error())
    return utils.take(R, (n, n, n, n))


def density(mol, imax: int = 200, conv: float = 1e-6) -> np.ndarray:
    atm, bas, env, nelec = utils.prep(mol)
    n = utils.angl(bas, 0)
    P = library.density_c(*utils.mol_args(atm, bas, env), nelec, imax, conv)
    if not P:
        raise RuntimeError(
            "librint SCF failed to produce a valid density (see stderr for the "
            "reason: non-convergence, tr(PS) != nelec, or PSP != 2P)")
    return utils.take(P, (n, n))


def energy(mol, P: np.ndarray) -> float:
    atm, bas, env, nelec = utils.prep(mol)
    P = np.ascontiguousarray(P, dtype=np.float64)
    return library.energy_c(*utils.mol_args(atm, bas, env),
                            P.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), P.size)


def scf(mol, imax: int = 200, conv: float = 1e-6) -> float:
    atm, bas, env, nelec = utils.prep(mol)
    E = library.scf_c(*utils.mol_args(atm, bas, env), nelec, imax, conv)
    if np.isnan(E):
        raise RuntimeError(
            "librint SCF failed (see stderr for the reason: non-convergence, "
            "tr(PS) != nelec, or PSP != 2P)")
    return E


def _coord_code(coord):
    c = {'cart': 0, 'sph': 1}.get(coord)
    if c is None:
        raise ValueError("coordinate type does not exist: cart, sph")
    return c


def fock(mol, P: np.ndarray, schwarz_tol: float = 0.0, coord: str = 'cart') -> np.ndarray:
    """Stateless G = J - K/2 (the GPU integral2e_fock, reference src/scf.rs:316-410);
    coord 'cart' or 'sph' (P and G in that basis)."""
    atm, bas, env, nelec = utils.prep(mol)
    c = _coord_code(coord)
    n = utils.angl(bas, c)
    P = np.ascontiguousarray(P, dtype=np.float64)
    G = library.fock_c(*utils.mol_args(atm, bas, env), P.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                       P.size, c, 0, schwarz_tol)
    if not G:
        raise RuntimeError("librint fock failed: " + utils.last_error())
    return utils.take(G, (n, n))
