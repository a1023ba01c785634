"""Array plumbing for the C ABI (reference python/librint/utils.py:9-90)."""
import ctypes

import numpy as np


def _lib():
    from ._bindings import library
    return library


def take(ptr, shape):
    """Copy a returned buffer into numpy, then release it with free_c."""
    shape = tuple(int(s) for s in shape)
    out = np.array(np.ctypeslib.as_array(ptr, shape=shape), copy=True)
    _lib().free_c(ptr, int(np.prod(shape)))
    return out


def prep(mol):
    atm = np.ascontiguousarray(mol._atm, dtype=np.int32)
    bas = np.ascontiguousarray(mol._bas, dtype=np.int32)
    env = np.ascontiguousarray(mol._env, dtype=np.double)
    nelec = mol.nelec[0] + mol.nelec[1]
    return atm, bas, env, nelec


def angl(bas, coord):
    n = 0
    for b in np.asarray(bas).reshape(-1, 8):
        l, nctr = int(b[1]), int(b[3])
        n += ((l + 1) * (l + 2) // 2 if coord == 0 else 2 * l + 1) * nctr
    return n


def mol_args(atm, bas, env):
    """(ptr, len) triplets for the libcint arrays."""
    return (atm.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), atm.size,
            bas.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), bas.size,
            env.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), env.size)


def last_error():
    return _lib().librint_last_error().decode()
