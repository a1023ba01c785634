"""The oracle (oracle/refcpu.c) against every pin the reference holds for
this path, plus mathematical pins for what those leave open."""
import itertools
import json
import os

import numpy as np
import pytest

from conftest import GOLD, ref_mol, ref_truth
from oracle import refcpu
import textbook_eri as tb

G = json.load(open(os.path.join(GOLD, "rys_mpmath.json")))


@pytest.mark.parametrize("name,nelec,n", [("h2", 2, 2), ("h2o", 10, 7)])
@pytest.mark.parametrize("typ,key", [(0, "S"), (1, "T"), (2, "V")])
def test_int1e_truth_sto3g(name, nelec, n, typ, key):
    m = ref_mol(name, "sto3g", nelec)
    M = refcpu.int1e(m, typ)
    ref = ref_truth(name, "sto3g_" + key).reshape(n, n)
    assert np.abs(M - ref).max() < 1e-12  # reference checks/overlap.rs:20 epsilon


@pytest.mark.parametrize("typ,key", [(0, "S"), (1, "T"), (2, "V")])
def test_int1e_truth_c6h6_631g_dshells(typ, key):
    m = ref_mol("c6h6", "631g", 42)
    M = refcpu.int1e(m, typ)
    ref = ref_truth("c6h6", "631g_" + key).reshape(102, 102)
    assert np.abs(M - ref).max() < 1e-12


@pytest.mark.parametrize("name,nelec", [("h2", 2), ("h2o", 10)])
def test_scf_energy_truth(name, nelec):
    # reference checks/energy.rs:182-196: imax 200, conv 1e-6 on |dP|/2
    m = ref_mol(name, "sto3g", nelec)
    E = refcpu.scf(m, 200, 1e-6)
    Eref = float(ref_truth(name, "sto3g_E"))
    assert abs(E - Eref) < 1e-10


def test_boys_vs_mpmath():
    for e in G["boys"]:
        F = refcpu.boys(18, e["x"])
        ref = np.array(e["F"])
        assert np.all(np.abs(F - ref) <= 4e-16 * np.abs(ref) + 1e-300), e["x"]


def test_rys_vs_mpmath():
    for e in G["rys"]:
        v, w = refcpu.rys_nodes(e["n"], e["x"])
        assert np.abs(v - e["v"]).max() < 5e-16, (e["n"], e["x"])
        assert np.abs(w - e["w"]).max() < 5e-16, (e["n"], e["x"])


@pytest.mark.parametrize("n", [1, 3, 5, 9, 10])
def test_rys_moments(n):
    # sum_r w_r v_r^k = F_k(x) for k < 2n (Gauss rule exactness)
    for x in [0.0, 0.3, 2.0, 9.0, 27.0, 44.0, 70.0, 120.0]:
        v, w = refcpu.rys_nodes(n, x)
        F = refcpu.boys(2 * n - 1, x)
        for k in range(2 * n):
            assert abs(np.sum(w * v ** k) - F[k]) <= 1e-15 * F[0] + 1e-14 * F[k], (n, x, k)


def test_rys_large_x_hermite_limit():
    # t_r^2 -> h_r / x and w_r -> H_r / sqrt(x) with the Gauss-Hermite rule
    for n in (1, 2, 4):
        s, h = np.polynomial.hermite.hermgauss(2 * n)
        pos = s > 0
        x = 400.0
        v, w = refcpu.rys_nodes(n, x)
        assert np.allclose(v, np.sort(s[pos] ** 2) / x, rtol=1e-13, atol=0)
        assert np.allclose(w, h[pos][np.argsort(s[pos] ** 2)] / np.sqrt(x), rtol=1e-13, atol=0)


def test_ssss_closed_form():
    # (ss|ss) for single primitives: 2 pi^{5/2}/(p q sqrt(p+q)) e^{-..} e^{-..} F0(T)
    from librint_amd import mol
    atoms = [(1, (0.0, 0.0, 0.0)), (1, (0.0, 0.3, 1.1)), (1, (0.7, -0.2, 2.0)), (1, (1.5, 0.4, -0.3))]
    basis = {1: [(0, [0.9], [1.0])]}
    m = mol.build(atoms, basis, normalized=True)
    c = 1.0 * 0.282094791773878143
    A = np.array([a[1] for a in atoms])
    a = 0.9
    p = q = 2 * a
    P = (A[0] + A[1]) / 2
    Q = (A[2] + A[3]) / 2
    T = p * q / (p + q) * np.sum((P - Q) ** 2)
    ref = 2 * np.pi ** 2.5 / (p * q * np.sqrt(p + q)) * np.exp(-a / 2 * np.sum((A[0] - A[1]) ** 2)) \
        * np.exp(-a / 2 * np.sum((A[2] - A[3]) ** 2)) * tb.boys(0, T)[0] * c ** 4
    got = refcpu.int2e_shell(m, (0, 1, 2, 3))[0, 0, 0, 0]
    assert abs(got - ref) < 1e-15


def test_eri_vs_textbook_h2o_all_quartets():
    m = ref_mol("h2o", "sto3g", 10)
    for q in itertools.product(range(m.nbas), repeat=4):
        assert np.abs(tb.eri_block(m, q) - refcpu.int2e_shell(m, q)).max() < 1e-13, q


def test_eri_vs_textbook_d_shells():
    m = ref_mol("h2o", "def2svp", 10)
    rng = np.random.default_rng(7)
    qs = [(5, 5, 5, 5), (5, 3, 5, 0), (11, 5, 2, 5)] + [tuple(int(v) for v in rng.integers(0, m.nbas, 4))
                                                        for _ in range(10)]
    for q in qs:
        assert np.abs(tb.eri_block(m, q) - refcpu.int2e_shell(m, q)).max() < 1e-13, q


def test_eri_symmetry_and_schwarz():
    m = ref_mol("h2o", "sto3g", 10)
    R = refcpu.int2e(m)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.abs(R - R.transpose(perm)).max() < 1e-14
    n = R.shape[0]
    M = R.reshape(n * n, n * n)
    assert np.linalg.eigvalsh(M).min() > -1e-12  # (ij|kl) as a pair matrix is PSD
    d = np.sqrt(np.abs(np.diag(M)))
    assert np.all(np.abs(M) <= np.outer(d, d) + 1e-14)


def test_fock_direct_matches_tensor():
    # integral2e_fock (src/scf.rs:316-410) == calc_F - H (src/scf.rs:651-670)
    m = ref_mol("h2o", "sto3g", 10)
    rng = np.random.default_rng(3)
    P = rng.normal(size=(7, 7))
    P = P + P.T
    R = refcpu.int2e(m)
    J = np.einsum("ijkl,lk->ij", R, P)
    K = np.einsum("ijkl,jk->il", R, P)
    G = refcpu.fock2e(m, P)
    assert np.abs(G - (J - 0.5 * K)).max() < 1e-13


def test_screened_symmetric_fock_matches_reference_loop():
    # the threaded 8-fold-symmetry comparator against the reference-loop
    # fixture (every ordered quartet, src/scf.rs:316-410): same G at tau = 0,
    # partitions of the bra pairs add up, thread count does not matter
    from librint_amd import mol
    d = np.load(os.path.join(GOLD, "fock_c2h6_631g.npz"))
    m = mol.alkane_631g(2)
    Q = refcpu.schwarz(m, 4)
    G, cnt = refcpu.fock2e_screened(m, d["P"], 0.0, Q=Q, nthreads=4)
    assert np.abs(G - d["G"]).max() < 1e-13 * np.abs(d["G"]).max()
    nz = np.count_nonzero(Q)
    assert cnt == nz * (nz + 1) // 2
    G1, _ = refcpu.fock2e_screened(m, d["P"], 0.0, Q=Q, nthreads=1)
    assert np.abs(G1 - G).max() < 1e-14 * np.abs(G).max()
    parts = [refcpu.fock2e_screened(m, d["P"], 0.0, Q=Q, stride=3, phase=k, nthreads=2) for k in range(3)]
    assert sum(c for _, c in parts) == cnt
    assert np.abs(sum(g for g, _ in parts) - G).max() < 1e-14 * np.abs(G).max()
    # a screened build: fewer quartets, error within tau * sum|P| per element
    Gs, cs = refcpu.fock2e_screened(m, d["P"], 1e-8, Q=Q, nthreads=4)
    assert cs < cnt
    assert np.abs(Gs - d["G"]).max() < 1e-8 * np.abs(d["P"]).sum()


def test_planted_mistake_is_caught():
    # a wrong sign in the exchange term must fail the tensor comparison above
    m = ref_mol("h2o", "sto3g", 10)
    rng = np.random.default_rng(3)
    P = rng.normal(size=(7, 7))
    P = P + P.T
    R = refcpu.int2e(m)
    wrong = np.einsum("ijkl,lk->ij", R, P) + 0.5 * np.einsum("ijkl,jk->il", R, P)
    assert np.abs(refcpu.fock2e(m, P) - wrong).max() > 1e-3


def test_eri_vs_textbook_f_g_shells():
    """l = 3, 4 quartets of the oracle against the McMurchie-Davidson
    restatement (f on O contracted from two primitives, g uncontracted)."""
    from librint_amd import mol
    m = mol.h2o_fg()
    # shells: O 0 s, 1 s, 2 p, 3 d, 4 f, 5 g; H 6 s, 7 p, 8 d; H 9 s, 10 p, 11 d
    for q in [(5, 0, 0, 0), (4, 4, 0, 0), (5, 2, 3, 6), (4, 3, 4, 9), (5, 5, 0, 1), (4, 8, 2, 11),
              (5, 4, 7, 0), (0, 5, 6, 4), (10, 4, 5, 9)]:
        a = tb.eri_block(m, q)
        assert np.abs(a - refcpu.int2e_shell(m, q)).max() < 1e-14 * max(1.0, np.abs(a).max()), q


@pytest.mark.parametrize("typ", [0, 1, 2])
def test_int1e_vs_textbook_f_g_shells(typ):
    """Every one-electron block of the f/g molecule: oracle vs the
    McMurchie-Davidson S, T (ket Laplacian), V (Hermite R integrals)."""
    from librint_amd import mol
    m = mol.h2o_fg()
    R = refcpu.int1e(m, typ)
    off = np.cumsum([0] + [(int(b[1]) + 1) * (int(b[1]) + 2) // 2 for b in m._bas])
    for i in range(m.nbas):
        for j in range(m.nbas):
            blk = tb.int1e_block(m, i, j, typ)
            assert np.abs(blk - R[off[i]:off[i + 1], off[j]:off[j + 1]]).max() < 1e-13, (typ, i, j)


def test_eri_vs_textbook_general_contraction():
    """The oracle's nctr > 1 branch (hierarchical primitive -> contraction
    loop of reference src/eri.rs:571-692) against the McMurchie-Davidson
    restatement contracting each column separately, on H2O with pairs of
    contractions over shared s and p primitives."""
    from librint_amd import mol
    m = mol.h2o_gc()
    assert max(int(b[3]) for b in m._bas) == 2
    # shells: O 0 s(x2), 1 p(x2), 2 d; H 3 s(x2), 4 p; H 5 s(x2), 6 p
    for q in [(0, 0, 0, 0), (0, 1, 2, 3), (1, 1, 1, 1), (2, 0, 3, 0), (4, 5, 0, 1), (1, 3, 5, 2),
              (6, 0, 1, 5), (3, 3, 5, 5)]:
        a = tb.eri_block(m, q)
        assert np.abs(a - refcpu.int2e_shell(m, q)).max() < 1e-13 * max(1.0, np.abs(a).max()), q


@pytest.mark.parametrize("typ", [0, 1, 2])
def test_int1e_vs_textbook_general_contraction(typ):
    from librint_amd import mol
    m = mol.h2o_gc()
    R = refcpu.int1e(m, typ)
    off = np.cumsum([0] + [(int(b[1]) + 1) * (int(b[1]) + 2) // 2 * int(b[3]) for b in m._bas])
    assert off[-1] == m.nao() == 24
    for i in range(m.nbas):
        for j in range(m.nbas):
            blk = tb.int1e_block(m, i, j, typ)
            assert np.abs(blk - R[off[i]:off[i + 1], off[j]:off[j + 1]]).max() < 1e-13, (typ, i, j)


def _charged_mol():
    """H2O with a third, negative point charge carrying an H basis: the
    nuclear repulsion uses the signed charges (reference src/scf.rs:946-953),
    the attraction |Z| (reference src/g1e.rs:270-274)."""
    from librint_amd import mol
    atoms = [(8, (0.0, 0.0, 0.0)), (1, (1.43, 1.11, 0.0)), (-1, (-1.6, 0.9, 0.7))]
    return mol.build(atoms, {8: mol.STO3G[8], 1: mol.STO3G[1], -1: mol.STO3G[1]}, "h2o_neg", normalized=True)


def test_nuclear_repulsion_signed_charges():
    m = _charged_mol()
    R = m.coords()
    Z = m.charges().astype(float)
    ref = sum(Z[i] * Z[j] / np.linalg.norm(R[i] - R[j]) for i in range(3) for j in range(i))
    absz = sum(abs(Z[i] * Z[j]) / np.linalg.norm(R[i] - R[j]) for i in range(3) for j in range(i))
    assert abs(ref - absz) > 1.0  # |Z| in the repulsion would be caught
    E = refcpu.energy(m, np.zeros((m.nao(), m.nao())))
    assert abs(E - ref) < 1e-13
