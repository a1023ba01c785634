"""Host side of the learned SCF guess (no GPU): the covariant block model
round-trips, features are rotation invariant, the geometry sampler is seeded
and clash-free."""
import numpy as np
import pytest


def _sample(nc, seed):
    from librint_amd import guess as G
    return G.alkane_sample(nc, np.random.default_rng(seed))


@pytest.mark.parametrize("nc", [2, 4, 6])
def test_block_model_round_trip(nc):
    from librint_amd import guess as G
    m = _sample(nc, nc)
    si = G.ShellInfo(m)
    X, geo = G.features(m, si)
    A, B, u = geo
    assert X.shape == (len(A), G.NFEAT)
    assert len(A) == len(si.l) * (len(si.l) + 1) // 2
    t = np.random.default_rng(1).normal(size=(len(A), 2))
    pp = (si.l[A] == 1) & (si.l[B] == 1)
    t[~pp, 1] = 0.0
    P = G.assemble(m, t, si, geo)
    assert P.shape == (m.nao(), m.nao())
    np.testing.assert_array_equal(P, P.T)
    T, W = G.targets(m, P, si, geo)
    np.testing.assert_allclose(T * W, t * W, atol=2e-6)
    assert np.all(W[pp] == 1.0) and np.all(W[~pp, 1] == 0.0)


def test_features_rotation_invariant_and_blocks_covariant():
    from librint_amd import guess as G, mol as M
    m = _sample(5, 7)
    X, geo = G.features(m)
    R = M.random_rotation(np.random.default_rng(3))
    env = m._env.copy()
    for a in range(len(m._atm)):
        p = int(m._atm[a][1])
        env[p:p + 3] = R @ m._env[p:p + 3]
    m2 = M.Mole(m._atm.copy(), m._bas.copy(), env, m.nelectron)
    X2, geo2 = G.features(m2)
    np.testing.assert_allclose(X2, X, atol=1e-6)
    # the same coefficients give the rotated density: p blocks transform with R
    t = np.random.default_rng(2).normal(size=(X.shape[0], 2))
    P = G.assemble(m, t)
    P2 = G.assemble(m2, t)
    si = G.ShellInfo(m)
    U = np.eye(m.nao())
    for s in range(len(si.l)):
        if si.l[s] == 1:
            a = si.ao[s]
            U[a:a + 3, a:a + 3] = R
    np.testing.assert_allclose(P2, U @ P @ U.T, atol=1e-12)


def test_sampler_seeded_and_clash_free():
    from librint_amd import guess as G
    a = G.perturbed_alkane(6, np.random.default_rng(5))
    b = G.perturbed_alkane(6, np.random.default_rng(5))
    np.testing.assert_array_equal(np.array([x for _, x in a]), np.array([x for _, x in b]))
    X = np.array([x for _, x in a])
    Z = np.array([z for z, _ in a])
    d = np.linalg.norm(X[:, None] - X[None], axis=-1)
    np.fill_diagonal(d, 10.0)
    cc = d[np.ix_(Z == 6, Z == 6)]
    assert np.all(np.sort(cc, axis=1)[:, 0] > 1.3) and np.all(np.sort(cc, axis=1)[:, 0] < 1.8)
    ch = d[np.ix_(Z == 1, Z == 6)].min(axis=1)
    assert np.all((ch > 0.85) & (ch < 1.35))
    assert d[np.ix_(Z == 1, Z == 1)].min() > 1.4


def test_rejects_basis_outside_the_model():
    from librint_amd import guess as G, mol as M
    with pytest.raises(ValueError):
        G.ShellInfo(M.h2o_631gstar_like())
