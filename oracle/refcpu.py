"""ctypes wrapper over oracle/_build/librefcpu.so -- test infrastructure only.

Only tests/, the smoke check and the CPU-baseline leg of the benchmark import
this module.  Every function here runs the plain-C restatement of the
reference algorithm in refcpu.c on the host.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "librefcpu.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        _lib = ctypes.CDLL(SO)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        _lib.rc_boys.argtypes = (ctypes.c_int, ctypes.c_double, dp)
        _lib.rc_rys_roots.argtypes = (ctypes.c_int, ctypes.c_double, dp, dp)
        _lib.rc_rys_roots_v.argtypes = (ctypes.c_int, ctypes.c_double, dp, dp)
        _lib.rc_int2e_shell.argtypes = (dp, ip, ip, ctypes.c_int, ip, ctypes.c_int, dp)
        _lib.rc_int2e_shell.restype = ctypes.c_int
        _lib.rc_int2e.argtypes = (dp, ip, ctypes.c_int, ip, ctypes.c_int, dp)
        _lib.rc_fock2e.argtypes = (dp, dp, ip, ctypes.c_int, ip, ctypes.c_int, dp)
        _lib.rc_fock2e_sample.argtypes = (dp, dp, ip, ctypes.c_long, ip, ctypes.c_int, ip, ctypes.c_int, dp)
        _lib.rc_calc_F.argtypes = (ctypes.c_int, dp, dp, dp, dp)
        _lib.rc_int1e.argtypes = (dp, ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int, dp)
        _lib.rc_energyfast.argtypes = (dp, ip, ctypes.c_int, ip, ctypes.c_int, dp)
        _lib.rc_energyfast.restype = ctypes.c_double
        _lib.rc_density.argtypes = (dp, ip, ip, ctypes.c_int, ip, ctypes.c_int, dp, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_int)
        _lib.rc_density.restype = ctypes.c_int
        _lib.rc_scf.argtypes = (ip, ctypes.c_int, ip, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.c_int)
        _lib.rc_scf.restype = ctypes.c_double
        _lib.rc_nao.argtypes = (ip, ctypes.c_int, ctypes.c_int)
        _lib.rc_schwarz.argtypes = (dp, ip, ctypes.c_int, ip, ctypes.c_int, dp, ctypes.c_int)
        _lib.rc_fock2e_screened.argtypes = (dp, dp, dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ip,
                                            ctypes.c_int, ip, ctypes.c_int, dp, ctypes.c_int)
        _lib.rc_fock2e_screened.restype = ctypes.c_long
        _lib.rc_nao.restype = ctypes.c_int
    return _lib


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _mol(m):
    atm, bas, env = m.arrays()
    return atm, bas, env, len(atm) // 6, len(bas) // 8


def boys(mmax, x):
    out = np.zeros(mmax + 1)
    lib().rc_boys(mmax, x, _d(out))
    return out


def rys_roots(n, x):
    """(u, w) with u = t^2/(1-t^2), the reference's output convention."""
    u = np.zeros(n)
    w = np.zeros(n)
    lib().rc_rys_roots(n, x, _d(u), _d(w))
    return u, w


def rys_nodes(n, x):
    """(v, w) with v = t^2."""
    v = np.zeros(n)
    w = np.zeros(n)
    lib().rc_rys_roots_v(n, x, _d(v), _d(w))
    return v, w


def int2e_shell(m, shls):
    atm, bas, env, natm, nbas = _mol(m)
    b = bas.reshape(-1, 8)
    dims = [((b[s, 1] + 1) * (b[s, 1] + 2) // 2) * b[s, 3] for s in shls]
    out = np.zeros(int(np.prod(dims)))
    sh = np.asarray(shls, dtype=np.int32)
    lib().rc_int2e_shell(_d(out), _i(sh), _i(atm), natm, _i(bas), nbas, _d(env))
    # buffer is i fastest -> (l,k,j,i) in C order
    return out.reshape(dims[::-1]).transpose(3, 2, 1, 0).copy()


def int2e(m):
    atm, bas, env, natm, nbas = _mol(m)
    n = m.nao()
    R = np.zeros(n ** 4)
    lib().rc_int2e(_d(R), _i(atm), natm, _i(bas), nbas, _d(env))
    return R.reshape(n, n, n, n)


def fock2e(m, P):
    atm, bas, env, natm, nbas = _mol(m)
    n = m.nao()
    P = np.ascontiguousarray(P, dtype=np.float64)
    G = np.zeros((n, n))
    lib().rc_fock2e(_d(G), _d(P), _i(atm), natm, _i(bas), nbas, _d(env))
    return G


def fock2e_sample(m, P, quartets):
    atm, bas, env, natm, nbas = _mol(m)
    n = m.nao()
    P = np.ascontiguousarray(P, dtype=np.float64)
    q = np.ascontiguousarray(quartets, dtype=np.int32)
    G = np.zeros((n, n))
    lib().rc_fock2e_sample(_d(G), _d(P), _i(q), len(q), _i(atm), natm, _i(bas), nbas, _d(env))
    return G


def fock2e_sample_threaded(m, P, quartets, nthreads=None):
    """fock2e_sample over the quartet list split into chunks, one host
    thread each (the C call releases the interpreter lock); partial G's
    summed."""
    from concurrent.futures import ThreadPoolExecutor
    q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
    nt = max(1, min(nthreads or host_threads(), len(q)))
    chunks = np.array_split(q, nt)
    with ThreadPoolExecutor(max_workers=nt) as ex:
        parts = list(ex.map(lambda c: fock2e_sample(m, P, c), chunks))
    return np.sum(parts, axis=0)


def host_threads():
    """Threads for the threaded comparator: the cores this process may use,
    at most OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def schwarz(m, nthreads=None):
    """Q[ij] = sqrt(max|(ij|ij)|) per shell pair, ij = i(i+1)/2 + j."""
    atm, bas, env, natm, nbas = _mol(m)
    Q = np.zeros(nbas * (nbas + 1) // 2)
    lib().rc_schwarz(_d(Q), _i(atm), natm, _i(bas), nbas, _d(env), nthreads or host_threads())
    return Q


def fock2e_screened(m, P, tau, Q=None, stride=1, phase=0, nthreads=None):
    """G = J - K/2 from the canonical quartets with Q_ij Q_kl > tau (8-fold
    symmetry, threads); returns (G, quartets evaluated).  stride/phase visit
    the bra pairs ij with ij % stride == phase only."""
    atm, bas, env, natm, nbas = _mol(m)
    n = m.nao()
    if Q is None:
        Q = schwarz(m, nthreads)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    P = np.ascontiguousarray(P, dtype=np.float64)
    G = np.zeros((n, n))
    cnt = lib().rc_fock2e_screened(_d(G), _d(P), _d(Q), float(tau), int(stride), int(phase), _i(atm), natm,
                                   _i(bas), nbas, _d(env), nthreads or host_threads())
    return G, int(cnt)


def int1e(m, typ):
    atm, bas, env, natm, nbas = _mol(m)
    n = m.nao()
    R = np.zeros((n, n))
    lib().rc_int1e(_d(R), int(typ), _i(atm), natm, _i(bas), nbas, _d(env))
    return R


def energy(m, P):
    atm, bas, env, natm, nbas = _mol(m)
    P = np.ascontiguousarray(P, dtype=np.float64)
    return lib().rc_energyfast(_d(P), _i(atm), natm, _i(bas), nbas, _d(env))


def density(m, imax=200, conv=1e-6, store_tensor=True):
    atm, bas, env, natm, nbas = _mol(m)
    n = m.nao()
    P = np.zeros((n, n))
    it = np.zeros(1, dtype=np.int32)
    rc = lib().rc_density(_d(P), _i(it), _i(atm), natm, _i(bas), nbas, _d(env), m.nelectron, imax, conv,
                          int(store_tensor))
    return rc, P, int(it[0])


def scf(m, imax=200, conv=1e-6, store_tensor=True):
    atm, bas, env, natm, nbas = _mol(m)
    return lib().rc_scf(_i(atm), natm, _i(bas), nbas, _d(env), m.nelectron, imax, conv, int(store_tensor))
