"""The 32-water even-tempered cluster (512 shells, 992 cartesian functions,
s, p and d shells, up to 5 Rys roots) at full size on the HIP path, against
the oracle: shell blocks of every angular class, the whole Fock matrix for a
density supported on a few shells, and the mixed-precision energy error."""
import sys

import numpy as np
import pytest

from oracle import refcpu

pytestmark = pytest.mark.gpu

TAU = 1e-12


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def w32():
    from librint_amd import mol
    from librint_amd.context import FockContext
    m = mol.water_cluster(32)
    ctx = FockContext(m, schwarz_tol=TAU)
    yield m, ctx
    ctx.close()


def _core_density(m):
    from librint_amd import scf
    S = scf.int1e(m, "ovlp")
    H = scf.int1e(m, "kin") + scf.int1e(m, "nuc")
    s, U = np.linalg.eigh(S)
    X = U / np.sqrt(s)
    e, C = np.linalg.eigh(X.T @ H @ X)
    C = X @ C
    nocc = m.nelectron // 2
    return 2.0 * C[:, :nocc] @ C[:, :nocc].T


def test_w32_shell_blocks_every_class(w32):
    """Four seeded shell quartets of every (la lb|lc ld) class, each in all
    its orderings' canonical form, against the oracle's integral."""
    m, ctx = w32
    _say("w32 shell blocks of every class")
    ls = np.array([b[1] for b in m._bas])
    by_l = {l: np.flatnonzero(ls == l) for l in (0, 1, 2)}
    rng = np.random.default_rng(32)
    qs = []
    for la in range(3):
        for lb in range(la, 3):
            for lc in range(3):
                for ld in range(lc, 3):
                    for _ in range(4):
                        qs.append(tuple(int(rng.choice(by_l[x])) for x in (la, lb, lc, ld)))
    # near-diagonal quartets on one water (largest integrals)
    qs += [(0, 1, 2, 3), (7, 7, 7, 7), (5, 6, 5, 6), (8, 0, 7, 5)]
    got = ctx.shells(qs)
    worst = 0.0
    for q, blk in zip(qs, got):
        ref = refcpu.int2e_shell(m, q)
        worst = max(worst, np.abs(blk - ref).max())
    _say("quartets", len(qs), "max |diff|", worst)
    assert worst < 1e-12


def test_w32_fock_block_density(w32):
    """Full-size Fock build for a density supported on an s and a d shell of
    one water and a p shell of another: the oracle visits every ordered
    quartet (i j|k l) that reads it (Coulomb P_lk: (i j|a b); exchange
    P_jk: (i a|b l); a and b in the support) under the same Schwarz product
    screen."""
    m, ctx = w32
    _say("w32 Fock, block density")
    n = m.nao()
    nb = m.nbas
    bas = m._bas
    ao = np.concatenate([[0], np.cumsum([(b[1] + 1) * (b[1] + 2) // 2 for b in bas])])
    ls = [b[1] for b in bas]
    S = [1, 7, 17 + ls[16:].index(1)]  # O s (free), O d, a p of the second water
    assert [ls[s] for s in S] == [0, 2, 1]
    P = np.zeros((n, n))
    rng = np.random.default_rng(9)
    idx = np.concatenate([np.arange(ao[s], ao[s + 1]) for s in S])
    B = rng.normal(size=(len(idx), len(idx)))
    P[np.ix_(idx, idx)] = B + B.T
    G = ctx.fock(P)
    Qp = refcpu.schwarz(m)
    Q = np.zeros((nb, nb))
    iu = np.tril_indices(nb)
    Q[iu] = Qp
    Q = np.maximum(Q, Q.T)
    i, j = np.meshgrid(np.arange(nb), np.arange(nb), indexing="ij")
    i, j = i.ravel(), j.ravel()
    quartets = []
    for a in S:
        for b in S:
            Jq = np.stack([i, j, np.full_like(i, a), np.full_like(i, b)], 1)
            Jq = Jq[Q[i, j] * Q[a, b] > TAU]
            Kq = np.stack([i, np.full_like(i, a), np.full_like(i, b), j], 1)
            Kq = Kq[Q[i, a] * Q[b, j] > TAU]
            quartets += [Jq, Kq]
    q = np.unique(np.concatenate(quartets), axis=0).astype(np.int32)
    _say("oracle quartets", len(q))
    Gr = refcpu.fock2e_sample_threaded(m, P, q)
    scale = np.abs(Gr).max()
    _say("max |G - oracle|", np.abs(G - Gr).max(), "scale", scale)
    assert np.abs(G - Gr).max() < 1e-11 * scale
    assert np.abs(G - G.T).max() < 1e-13 * np.abs(G).max()


def test_w32_mixed_precision_energy(w32):
    """Mixed precision at full size: for every tau_mix in 1e-8..1e-5 some
    quartets run FP32, the energy 1/2 tr(P G) moves by at most 1 uEh from
    the FP64 build, and the FP32 share grows with tau_mix."""
    m, ctx = w32
    _say("w32 mixed precision")
    P = _core_density(m)
    G64 = ctx.fock(P)
    shares = []
    try:
        for tau in (1e-8, 1e-7, 1e-6, 1e-5):
            ctx.set_precision(1, tau)
            Gm = ctx.fock(P)
            fc = ctx.flag_counts()["quartets"]
            dE = 0.5 * np.sum(P * (Gm - G64))
            _say("tau_mix", tau, "fp64/fp32 quartets", fc[:2], "dE", dE, "max dG", np.abs(Gm - G64).max())
            assert fc[1] > 0 and fc[2] == 0
            assert sum(fc) == ctx.stats()["quartets"]
            assert abs(dE) <= 1e-6
            shares.append(fc[1])
    finally:
        ctx.set_precision(0, 0.0, 0.0)
    assert shares == sorted(shares)
    assert np.abs(ctx.fock(P) - G64).max() < 1e-12
