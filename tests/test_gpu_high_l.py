"""f and g shells (l = 3, 4) on the GPU against the oracle: the run-time
shaped two-electron kernels (tensor, shell blocks, Schwarz-screened Fock
build), the one-electron kernels, the spherical transform and the SCF."""
import sys

import numpy as np
import pytest

from oracle import refcpu

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rand_sym(n, seed):
    rng = np.random.default_rng(seed)
    P = rng.normal(size=(n, n))
    return 0.5 * (P + P.T)


@pytest.fixture(scope="module")
def mfg():
    from librint_amd import mol
    return mol.h2o_fg()


@pytest.fixture(scope="module")
def tensor_ref(mfg):
    return refcpu.int2e(mfg)


def test_int1e_f_g(gpu, mfg):
    from librint_amd import scf
    _say("int1e f/g")
    for k, typ in enumerate(("ovlp", "kin", "nuc")):
        R = scf.int1e(mfg, typ)
        assert np.abs(R - refcpu.int1e(mfg, k)).max() < 1e-12, typ


def test_int2e_tensor_f_g(gpu, mfg, tensor_ref):
    from librint_amd import scf
    _say("int2e tensor f/g")
    R = scf.int2e(mfg)
    assert np.abs(R - tensor_ref).max() < 1e-12
    for perm in [(1, 0, 2, 3), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.abs(R - R.transpose(perm)).max() == 0.0


def test_shell_blocks_f_g(gpu, mfg):
    from librint_amd.context import FockContext
    _say("shell blocks f/g")
    ctx = FockContext(mfg)
    qs = [(5, 5, 5, 5), (4, 5, 4, 5), (5, 0, 0, 0), (0, 5, 6, 4), (10, 4, 5, 9), (4, 8, 2, 11), (11, 5, 3, 4)]
    for q, blk in zip(qs, ctx.shells(qs)):
        assert np.abs(blk - refcpu.int2e_shell(mfg, q)).max() < 1e-12, q


def test_fock_f_g_vs_oracle(gpu, mfg, tensor_ref):
    """Unscreened and screened Fock builds; the screened one against the
    oracle within the screening bound."""
    from librint_amd import scf
    _say("fock f/g")
    n = mfg.nao()
    for seed in (1, 2):
        P = _rand_sym(n, seed)
        Gr = refcpu.fock2e(mfg, P)
        # the tensor contraction restated: J - K/2 from the exact tensor
        Gt = np.einsum("ijkl,kl->ij", tensor_ref, P) - 0.5 * np.einsum("ikjl,kl->ij", tensor_ref, P)
        assert np.abs(Gr - Gt).max() < 1e-11 * np.abs(Gt).max()
        G = scf.fock(mfg, P, 0.0)
        assert np.abs(G - Gr).max() < 1e-11 * max(1.0, np.abs(Gr).max())
        Gs = scf.fock(mfg, P, 1e-12)
        assert np.abs(Gs - Gr).max() < 1e-9 * max(1.0, np.abs(Gr).max())


def test_spherical_f_g(gpu, mfg):
    """coord = 1 with f and g shells: on-site overlap of each normalised
    single-primitive shell is the identity (the g shell; the f shell is a
    contraction of normalised primitives, so its block is a multiple of the
    identity), and the spherical tensor equals the cartesian one transformed
    by independently fitted solid harmonics."""
    from librint_amd import scf
    _say("spherical f/g")
    Ss = scf.int1e(mfg, "ovlp", "sph")
    # spherical order: O s s p(3) d(5) f(7) g(9) -> f at 10..16, g at 17..25
    np.testing.assert_allclose(Ss[17:26, 17:26], np.eye(9), atol=1e-12)
    f = Ss[10:17, 10:17]
    np.testing.assert_allclose(f, f[0, 0] * np.eye(7), atol=1e-12)
    ls = [int(b[1]) for b in mfg._bas]
    ncart, nsph = mfg.nao(0), mfg.nao(1)
    C = np.zeros((nsph, ncart))
    r = c = 0
    for l in ls:
        nc, ns = (l + 1) * (l + 2) // 2, 2 * l + 1
        C[r:r + ns, c:c + nc] = np.eye(1) if l == 0 else (np.eye(3) if l == 1 else _solid_harmonics(l))
        r += ns
        c += nc
    for typ in ("ovlp", "kin", "nuc"):
        np.testing.assert_allclose(scf.int1e(mfg, typ, "sph"), C @ scf.int1e(mfg, typ, "cart") @ C.T, atol=1e-12)
    Rc = scf.int2e(mfg, "cart")
    Rs = scf.int2e(mfg, "sph")
    ref = np.einsum("ai,bj,ck,dl,ijkl->abcd", C, C, C, C, Rc, optimize=True)
    assert np.abs(Rs - ref).max() < 1e-12


def _solid_harmonics(l):
    """cart -> real spherical rows of degree l fitted from scipy's spherical
    harmonics (r^l Y_lm on random points, least squares over the monomials
    in libcint order): the real combinations sqrt(2) (-1)^m Re/Im Y_l|m|,
    which cancel the Condon-Shortley phase."""
    import scipy.special as sp
    rng = np.random.default_rng(4 + l)
    v = rng.normal(size=(800, 3))
    x, y, z = v.T
    r = np.linalg.norm(v, axis=1)
    th, ph = np.arccos(z / r), np.arctan2(y, x)
    Y = lambda m: (sp.sph_harm_y(l, m, th, ph) if hasattr(sp, "sph_harm_y") else sp.sph_harm(m, l, ph, th))
    real = {0: Y(0).real}
    for m in range(1, l + 1):
        real[m] = np.sqrt(2) * (-1) ** m * Y(m).real
        real[-m] = np.sqrt(2) * (-1) ** m * Y(m).imag
    pows = [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]
    mono = np.stack([x ** a * y ** b * z ** c for a, b, c in pows], axis=1)
    rows = []
    for m in range(-l, l + 1):
        c, *_ = np.linalg.lstsq(mono, real[m] * r ** l, rcond=None)
        c[np.abs(c) < 1e-10] = 0.0
        rows.append(c)
    return np.array(rows)


def test_scf_f_g_vs_oracle(gpu, mfg):
    """Reference SCF protocol (P0 = I, Roothaan) with f and g shells: GPU
    energy equal to the oracle's within 1e-10 Eh."""
    from librint_amd import scf
    _say("scf f/g")
    E = scf.scf(mfg, imax=200, conv=1e-8)
    Er = refcpu.scf(mfg, imax=200, conv=1e-8)
    assert abs(E - Er) < 1e-10, (E, Er)


def test_fock_spherical_f_g(gpu, mfg):
    """fock_c with coord = 1 (reference integral2e_fock on the spherical
    basis) equals J - K/2 contracted from the spherical tensor int2e_c
    (coord = 1), which is pinned above."""
    from librint_amd import scf
    _say("spherical fock f/g")
    R = scf.int2e(mfg, 'sph')
    n = R.shape[0]
    rng = np.random.default_rng(21)
    P = rng.normal(size=(n, n))
    P = P + P.T
    G = scf.fock(mfg, P, 0.0, 'sph')
    Gr = np.einsum("ijkl,lk->ij", R, P) - 0.5 * np.einsum("ijkl,jk->il", R, P)
    assert G.shape == (n, n)
    assert np.abs(G - Gr).max() < 1e-11 * np.abs(Gr).max()

