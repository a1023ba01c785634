"""One-electron basis-parameter derivatives (GPU kernels): dHcore_c
(reference src/dscf.rs:320-336) and dS_u (src/dscf.rs:96-150) against
central finite differences of the oracle's one-electron integrals (an
independent restatement), and the env split rule (src/utils.rs:152-175)."""
import numpy as np
import pytest


def _mols():
    from librint_amd import mol
    return [mol.h2o_sto3g(), mol.h2o_631gstar_like(), mol.water_cluster(1)]


def test_split_rule():
    from librint_amd import dscf, mol
    m = mol.h2o_sto3g()
    s1, s2 = dscf.split(m._bas)
    bas = np.asarray(m._bas)
    assert s1 == bas[:, 5].min()
    # every exponent and coefficient of this basis lies in [s1, s2)
    for b in bas:
        assert s1 <= b[5] and b[5] + b[2] <= s2
        assert s1 <= b[6] and b[6] + b[2] * b[3] <= s2


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(3))
def test_dhcore_and_dsu_vs_fd(gpu, k):
    from librint_amd import dscf
    from oracle import refcpu
    m = _mols()[k]
    n = m.nao()
    rng = np.random.default_rng(1)
    P = rng.normal(size=(n, n))
    P = P + P.T
    s1, s2 = dscf.split(m._bas)
    dH = dscf.dHcoref(m, P)
    dSu = dscf.dSu(m)
    assert dSu.shape == (s2 - s1, n, n)
    env0 = m._env.copy()
    for j in range(s1, s2):
        h = 1e-5 * max(abs(env0[j]), 1e-2)
        vals, Ss = [], []
        for sg in (1, -1):
            m._env[j] = env0[j] + sg * h
            H = refcpu.int1e(m, 1) + refcpu.int1e(m, 2)
            vals.append(float((P * H).sum()))
            Ss.append(refcpu.int1e(m, 0))
            m._env[:] = env0
        fd = (vals[0] - vals[1]) / (2 * h)
        fdS = (Ss[0] - Ss[1]) / (2 * h)
        assert abs(fd - dH[j - s1]) < 1e-6 * max(1.0, abs(fd)), (j, fd, dH[j - s1])
        assert np.abs(fdS - dSu[j - s1]).max() < 1e-7, j
        # dS_u is symmetric in (a, b) like S itself
        np.testing.assert_allclose(dSu[j - s1], dSu[j - s1].T, atol=1e-15)


@pytest.mark.gpu
def test_dhcore_planted_mistake_is_caught(gpu):
    """A dropped exponent term (the -a_x G(a-1) / +2 shift pair) would show:
    the FD check distinguishes exponent derivatives from half of themselves."""
    from librint_amd import dscf
    from oracle import refcpu
    m = _mols()[1]
    n = m.nao()
    P = np.eye(n)
    s1, s2 = dscf.split(m._bas)
    dH = dscf.dHcoref(m, P)
    j = int(np.asarray(m._bas)[-1, 5])  # an exponent entry
    env0 = m._env.copy()
    h = 1e-5 * env0[j]
    m._env[j] = env0[j] + h
    ep = float((P * (refcpu.int1e(m, 1) + refcpu.int1e(m, 2))).sum())
    m._env[j] = env0[j] - h
    em = float((P * (refcpu.int1e(m, 1) + refcpu.int1e(m, 2))).sum())
    m._env[:] = env0
    fd = (ep - em) / (2 * h)
    assert abs(fd) > 1e-3
    assert abs(0.5 * dH[j - s1] - fd) > 1e-4
    assert abs(dH[j - s1] - fd) < 1e-6 * max(1.0, abs(fd))


def test_reference_gradient_entries_reject_bad_density():
    from librint_amd import dscf, mol
    m = mol.h2o_sto3g()
    with pytest.raises(RuntimeError):
        dscf.dHcoref(m, np.eye(3))
