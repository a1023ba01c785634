"""Persistent GPU context: basis staged once, Fock builds per iteration."""
import ctypes

import numpy as np

from ._bindings import library
from . import utils

STAT_KEYS = ("quartets", "prim_quartets", "integrals", "flops", "candidates", "launches",
             "kernel_ms", "schwarz_tol", "grad_kernel_ms", "scf_purified", "streams")


class FockContext:
    def __init__(self, mol, schwarz_tol=1e-14):
        atm, bas, env, nelec = utils.prep(mol)
        self._keep = (atm, bas, env)
        self.nelectron = nelec
        self.nbas = len(bas) // 8 if bas.ndim == 1 else bas.shape[0]
        self.h = library.librint_ctx_create(*utils.mol_args(atm, bas, env))
        if not self.h:
            raise RuntimeError("librint context creation failed: " + utils.last_error())
        self.nao = library.librint_ctx_nao(self.h)
        self.set_schwarz(schwarz_tol)

    def close(self):
        if getattr(self, "h", None):
            library.librint_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_shard(self, nshards, shard):
        """Keep one device's share of the Fock build (nshards devices; the
        shares' G's sum to the whole build)."""
        if library.librint_ctx_set_shard(self.h, int(nshards), int(shard)) != 0:
            raise RuntimeError(utils.last_error())

    def set_schwarz(self, tol):
        if library.librint_ctx_set_schwarz(self.h, float(tol)) != 0:
            raise RuntimeError(utils.last_error())

    def set_precision(self, mode, tau_mix=0.0, tau_skip=0.0):
        """mode 0 FP64; mode 1 mixed: per build, the 64-ket tasks whose
        largest Q_ab Q_cd max|D_block| is below tau_mix run FP32 recurrences;
        tau_skip > 0 drops the tasks below it (both modes)."""
        if library.librint_ctx_set_precision(self.h, int(mode), float(tau_mix), float(tau_skip), 0.0) != 0:
            raise RuntimeError(utils.last_error())

    def flag_counts(self):
        out = np.zeros(6)
        library.librint_ctx_flag_counts(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return {"quartets": out[:3].astype(int).tolist(), "prim_quartets": out[3:].astype(int).tolist()}

    def fock(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.shape == (self.nao, self.nao)
        G = np.empty_like(P)
        dp = ctypes.POINTER(ctypes.c_double)
        if library.librint_fock(self.h, P.ctypes.data_as(dp), G.ctypes.data_as(dp)) != 0:
            raise RuntimeError("librint fock failed: " + utils.last_error())
        return G

    def fock_device(self, P_dev, G_dev, stream=None):
        """P_dev, G_dev: contiguous float64 torch CUDA tensors (nao x nao)."""
        assert P_dev.is_contiguous() and G_dev.is_contiguous()
        st = stream.cuda_stream if stream is not None else None
        if library.librint_fock_device(self.h, ctypes.c_void_p(P_dev.data_ptr()),
                                       ctypes.c_void_p(G_dev.data_ptr()), ctypes.c_void_p(st)) != 0:
            raise RuntimeError("librint fock failed: " + utils.last_error())

    def shells(self, quartets):
        """List of (i,j,k,l) shell quartets -> list of (di,dj,dk,dl) blocks."""
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.int32).reshape(-1, 4))
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        n = library.librint_int2e_shells(self.h, q.ctypes.data_as(ip), len(q), None)
        if n == ctypes.c_size_t(-1).value:
            raise RuntimeError(utils.last_error())
        out = np.zeros(max(n, 1))
        library.librint_int2e_shells(self.h, q.ctypes.data_as(ip), len(q), out.ctypes.data_as(dp))
        bas = self._keep[1].reshape(-1, 8)
        res, off = [], 0
        for s in q:
            dims = [((bas[x, 1] + 1) * (bas[x, 1] + 2) // 2) * bas[x, 3] for x in s]
            sz = int(np.prod(dims))
            res.append(out[off:off + sz].reshape(dims[::-1]).transpose(3, 2, 1, 0).copy())
            off += sz
        return res

    def schwarz(self):
        out = np.zeros(self.nbas * (self.nbas + 1) // 2)
        library.librint_schwarz(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def stats(self):
        out = np.zeros(len(STAT_KEYS))
        library.librint_ctx_stats(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(STAT_KEYS))
        return dict(zip(STAT_KEYS, out.tolist()))

    def class_stats(self):
        """Per angular-class launch: counts, flops and the last build's device ms."""
        n = library.librint_ctx_class_stats(self.h, None, 0)
        out = np.zeros(14 * max(n, 1))
        library.librint_ctx_class_stats(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
        keys = ("la", "lb", "lc", "ld", "quartets", "prim_quartets", "flops", "lanes", "ms", "tasks", "mixed",
                "variant", "drow", "reduce_ms")
        rows = [dict(zip(keys, out[14 * k:14 * k + 14].tolist())) for k in range(n)]
        for r in rows:
            v = int(r["variant"])
            ls = "%d, %d, %d, %d" % (r["la"], r["lb"], r["lc"], r["ld"])
            if v == 2:
                r["kernel"] = "jk_team_kernel<%s>" % ls
            elif v == 1:
                r["kernel"] = "jk_group_kernel<%s, %s, %s>" % (ls, "true" if r["mixed"] else "false",
                                                              "true" if r["drow"] else "false")
            elif v == 0:
                r["kernel"] = "jk_kernel<%s, double>" % ls
            elif v == 3:
                r["kernel"] = "hl_kernel<3>"  # run-time class shape (an f or g shell)
            else:
                r["kernel"] = None
        return rows

    def scf(self, imax=200, conv=1e-6, guess=0, diis=False, P0=None, conv_e=0.0):
        """(E, P, iterations); guess 0 identity, 1 core Hamiltonian, 2 P0.
        conv_e > 0 adds |E_k - E_(k-1)| <= conv_e to the density criterion."""
        P = np.zeros((self.nao, self.nao)) if P0 is None else np.ascontiguousarray(P0, dtype=np.float64).copy()
        it = ctypes.c_int(0)
        E = library.librint_scf_ex(self.h, int(self.nelectron), int(imax), float(conv), float(conv_e), int(guess),
                                   int(bool(diis)), P.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   ctypes.byref(it))
        return E, P, it.value
