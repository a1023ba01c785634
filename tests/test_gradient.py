"""Analytic gradients (adjoint Rys kernel + host one-electron derivatives) vs
central finite differences, through the C ABI and the torch autograd
Function (run on an MI355X).

Two levels:
  * fixed density: d/dtheta of E2 = 1/2 tr(P G[P]) (the reference's dR_c,
    src/dscf.rs:370-524) and of tr(P H) for exponents, coefficients and
    nuclear coordinates, FD of the GPU Fock build at the same P;
  * converged RHF: the full gradient dH + dR - 1/2 dS (+ nuclear repulsion),
    FD of re-converged SCF energies (the reference's own check,
    python/tests/test_gradient_fd.py:21-38), here with a 5-point stencil,
    steps 1e-4 bohr and 1e-4*max(zeta, 1) in the exponents.
The reference's Enzyme gradients cannot be run here (nightly rustc with
Enzyme is not in the image), so FD is the pin.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _e2(mol, P):
    """1/2 tr(P G[P]) from the GPU Fock build without screening."""
    from librint_amd._bindings import library
    from librint_amd import utils
    atm, bas, env, _ = utils.prep(mol)
    P = np.ascontiguousarray(P, dtype=np.float64)
    G = library.fock_c(*utils.mol_args(atm, bas, env), P.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), P.size,
                       0, 0, 0.0)
    assert G, utils.last_error()
    n = P.shape[0]
    G = utils.take(G, (n, n))
    return 0.5 * float((P * G).sum())


def _e1(mol, P):
    from librint_amd import scf
    return float((P * (scf.int1e(mol, 'kin') + scf.int1e(mol, 'nuc'))).sum())


def _mols():
    from librint_amd import mol
    return {"h2o_d": mol.h2o_631gstar_like, "w2": lambda: mol.water_cluster(2), "h2o_fg": mol.h2o_fg}


def _density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) * 0.2
    return A + A.T


def _coord_slots(mol):
    return [(a, d, int(mol._atm[a][1]) + d) for a in range(len(mol._atm)) for d in range(3)]


@pytest.mark.parametrize("name", ["h2o_d", "w2", "h2o_fg"])
def test_dR_fixed_density_vs_fd(gpu, name):
    """Exponent and coefficient derivatives of E2 (dR_c) at a fixed density."""
    from librint_amd import dscf
    m = _mols()[name]()
    P = _density(m.nao(), 3)
    s1, s2 = dscf.split(m._bas)
    _say("dR_c", name, "nao", m.nao(), "params", s2 - s1)
    dR = dscf.dRf(m, P)
    env0 = m._env.copy()
    worst = 0.0
    for k in range(s1, s2):
        h = 1e-5 * max(abs(env0[k]), 1e-3)
        m._env[k] = env0[k] + h
        ep = _e2(m, P)
        m._env[k] = env0[k] - h
        em = _e2(m, P)
        m._env[:] = env0
        fd = (ep - em) / (2 * h)
        err = abs(fd - dR[k - s1]) / max(1.0, abs(fd))
        worst = max(worst, err)
        assert err < 1e-6, (k, fd, dR[k - s1])
    _say("  worst rel err %.2e" % worst)


@pytest.mark.parametrize("name", ["h2o_d", "w2", "h2o_fg"])
def test_nuclear_fixed_density_vs_fd(gpu, name):
    """Nuclear-coordinate derivatives of E2 (adjoint kernel) and of tr(P H)
    (host, including the point-charge positions) at a fixed density."""
    from librint_amd.context import FockContext
    from librint_amd import autograd
    m = _mols()[name]()
    P = _density(m.nao(), 5)
    ctx = FockContext(m, 0.0)
    _, _, parts = autograd.gradient(ctx, P, parts=True)
    ctx.close()
    _say("nuclear fixed-P", name, "natm", len(m._atm))
    env0 = m._env.copy()
    h = 1e-4
    w2 = w1 = 0.0
    for a, d, k in _coord_slots(m):
        m._env[k] = env0[k] + h
        e2p, e1p = _e2(m, P), _e1(m, P)
        m._env[k] = env0[k] - h
        e2m, e1m = _e2(m, P), _e1(m, P)
        m._env[:] = env0
        f2 = (e2p - e2m) / (2 * h)
        f1 = (e1p - e1m) / (2 * h)
        w2 = max(w2, abs(f2 - parts["gR"][a, d]))
        w1 = max(w1, abs(f1 - parts["gH"][a, d]))
        assert abs(f2 - parts["gR"][a, d]) < 1e-6 * max(1.0, abs(f2)), (a, d, f2, parts["gR"][a, d])
        assert abs(f1 - parts["gH"][a, d]) < 1e-6 * max(1.0, abs(f1)), (a, d, f1, parts["gH"][a, d])
    _say("  worst abs err 2e %.2e 1e %.2e" % (w2, w1))


def test_translational_invariance(gpu):
    """sum over atoms of every nuclear part vanishes (the 2e kernel's four
    centres, the 1e operator-centre term)."""
    from librint_amd.context import FockContext
    from librint_amd import autograd
    m = _mols()["w2"]()
    P = _density(m.nao(), 7)
    ctx = FockContext(m, 0.0)
    _, _, parts = autograd.gradient(ctx, P, parts=True)
    ctx.close()
    for k in ("gH", "gR", "gS", "gN"):
        s = parts[k].sum(axis=0)
        scale = max(1.0, np.abs(parts[k]).max())
        assert np.abs(s).max() < 1e-10 * scale, (k, s)


@pytest.mark.parametrize("name,lanes", [("w2", "1"), ("w2", "2"), ("h2o_fg", "2")])
def test_v_derivative_lane_shapes_agree(gpu, name, lanes, monkeypatch):
    """The nuclear-attraction derivative kernel with fewer lanes per shell
    pair than point charges (each lane then takes several charges and adds
    each charge's gradient per primitive pair) gives the same one-electron
    parts as with one charge per lane (the default up to 64 charges)."""
    from librint_amd.context import FockContext
    from librint_amd import autograd, dscf
    m = _mols()[name]()
    P = _density(m.nao(), 9)
    ctx = FockContext(m, 0.0)
    monkeypatch.delenv("LIBRINT_V_LANES", raising=False)
    _, _, ref = autograd.gradient(ctx, P, parts=True)
    dH0 = dscf.dHcoref(m, P)
    monkeypatch.setenv("LIBRINT_V_LANES", lanes)
    _say("V derivative lanes", name, lanes, "natm", len(m._atm))
    _, _, got = autograd.gradient(ctx, P, parts=True)
    dH1 = dscf.dHcoref(m, P)
    ctx.close()
    scale = max(1.0, np.abs(ref["gH"]).max())
    assert np.abs(got["gH"] - ref["gH"]).max() < 1e-11 * scale
    assert np.abs(dH1 - dH0).max() < 1e-11 * max(1.0, np.abs(dH0).max())


def _exp_step(alpha):
    """Exponent step 1e-4*max(zeta, 1) for the 5-point stencil: re-converged
    energies carry ~5e-12 Eh of noise (FP64 atomics in the Fock build, SCF
    stopping point), and the diffuse exponents of a cluster basis have large
    higher derivatives, so a 3-point quotient at 1e-5 is only good to ~1e-6."""
    return 1e-4 * max(float(alpha), 1.0)


def _fd5(f, x0, h):
    """5-point central difference of f at x0 (error O(h^4))."""
    return (8.0 * (f(x0 + h) - f(x0 - h)) - (f(x0 + 2 * h) - f(x0 - 2 * h))) / (12.0 * h)


def _rhf_energy(m, P0=None):
    from librint_amd import autograd
    E, P, it, ctx = autograd.rhf(m, conv=1e-10, P0=P0)
    ctx.close()
    return E


def _check_rhf_fd(m, coord_sel, exp_sel, tol):
    from librint_amd import autograd
    E0, P0, it, ctx = autograd.rhf(m, conv=1e-10)
    gn, de = autograd.gradient(ctx, P0)
    ctx.close()
    _say("  E0 %.12f iterations %d" % (E0, it))
    env0 = m._env.copy()

    def energy_at(k):
        def f(x):
            m._env[k] = x
            try:
                return _rhf_energy(m, P0)
            finally:
                m._env[:] = env0
        return f

    worst = 0.0
    for a, d, k in coord_sel:
        fd = _fd5(energy_at(k), env0[k], 1e-4)
        worst = max(worst, abs(fd - gn[a, d]))
        assert abs(fd - gn[a, d]) < tol, ("coord", a, d, fd, gn[a, d])
    for k in exp_sel:
        fd = _fd5(energy_at(k), env0[k], _exp_step(env0[k]))
        worst = max(worst, abs(fd - de[k]))
        assert abs(fd - de[k]) < tol, ("exp", k, fd, de[k])
    _say("  worst abs err %.2e" % worst)
    return gn, de


def test_rhf_gradient_vs_fd_water_dimer(gpu):
    """Converged-RHF gradient, every nuclear coordinate and exponent of (H2O)2."""
    from librint_amd import autograd
    m = _mols()["w2"]()
    _say("RHF gradient (H2O)2 nao", m.nao())
    _check_rhf_fd(m, _coord_slots(m), autograd.exponent_slots(m), 1e-6)


def test_rhf_gradient_vs_fd_f_g(gpu):
    """Converged-RHF gradient with f and g shells (a contracted f and a g on
    O, d on H; up to 10 Rys roots in the adjoint): every nuclear coordinate,
    exponent and contraction coefficient against 5-point FD (coordinates,
    exponents) and central FD (coefficients) of re-converged energies."""
    from librint_amd import autograd, dscf
    m = _mols()["h2o_fg"]()
    _say("RHF gradient H2O f/g nao", m.nao())
    _check_rhf_fd(m, _coord_slots(m), autograd.exponent_slots(m), 1e-6)
    E0, P0, it, ctx = autograd.rhf(m, conv=1e-10)
    ctx.close()
    s1, s2 = dscf.split(m._bas)
    da = dscf.danalyticalf(m, P0)
    exps = set(autograd.exponent_slots(m).tolist())
    env0 = m._env.copy()
    worst, n = 0.0, 0
    for k in range(s1, s2):
        if k in exps:
            continue
        h = 1e-5 * max(abs(env0[k]), 1e-2)
        m._env[k] = env0[k] + h
        ep = _rhf_energy(m, P0)
        m._env[k] = env0[k] - h
        em = _rhf_energy(m, P0)
        m._env[:] = env0
        fd = (ep - em) / (2 * h)
        worst = max(worst, abs(fd - da[k - s1]))
        assert abs(fd - da[k - s1]) < 1e-6, (k, fd, da[k - s1])
        n += 1
    assert n > 0
    _say("  coefficients: %d, worst abs err %.2e" % (n, worst))


def test_rhf_gradient_reference_entries_agree(gpu):
    """danalytical_c / denergy_c (stateless, unscreened) equal the context
    gradient on env[s1:s2]; grad_c = dHcore + dR; dS_c consistent."""
    from librint_amd import autograd, dscf
    m = _mols()["h2o_d"]()
    E0, P, it, ctx = autograd.rhf(m, conv=1e-10)
    gn, de, parts = autograd.gradient(ctx, P, parts=True)
    ctx.close()
    s1, s2 = dscf.split(m._bas)
    da = dscf.danalyticalf(m, P)
    np.testing.assert_allclose(da, de[s1:s2], rtol=1e-10, atol=1e-11)
    # same expression; FP64 atomics make the kernel sums order-dependent at 1e-15
    np.testing.assert_allclose(dscf.denergyf(m, P), da, rtol=1e-12, atol=1e-13)
    g = dscf.grad(m, P)
    np.testing.assert_allclose(g, parts["dH"][s1:s2] + parts["dR"][s1:s2], rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(dscf.dSf(m, P), parts["dS"][s1:s2], rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(dscf.dHcoref(m, P), parts["dH"][s1:s2], rtol=1e-10, atol=1e-11)
    gnuc = dscf.grad_nuc(m, P)
    np.testing.assert_allclose(gnuc, gn, rtol=1e-9, atol=1e-11)
    # the Pulay term from dS_u: sum_ab (PFP)_ab dS_ab/denv
    from librint_amd import scf
    from librint_amd.context import FockContext
    c2 = FockContext(m, 0.0)
    F = scf.int1e(m, 'kin') + scf.int1e(m, 'nuc') + c2.fock(P)
    c2.close()
    Q = P @ F @ P
    dSu = dscf.dSu(m)
    np.testing.assert_allclose(np.einsum('kab,ab->k', dSu, Q), parts["dS"][s1:s2], rtol=1e-9, atol=1e-10)


def test_rhf_gradient_coefficients_vs_fd(gpu):
    """Coefficient entries of danalytical_c vs FD of re-converged energies."""
    from librint_amd import autograd, dscf
    m = _mols()["h2o_d"]()
    E0, P0, it, ctx = autograd.rhf(m, conv=1e-10)
    ctx.close()
    s1, s2 = dscf.split(m._bas)
    da = dscf.danalyticalf(m, P0)
    exps = set(autograd.exponent_slots(m).tolist())
    env0 = m._env.copy()
    n = 0
    for k in range(s1, s2):
        if k in exps:
            continue
        h = 1e-5 * max(abs(env0[k]), 1e-2)
        m._env[k] = env0[k] + h
        ep = _rhf_energy(m, P0)
        m._env[k] = env0[k] - h
        em = _rhf_energy(m, P0)
        m._env[:] = env0
        fd = (ep - em) / (2 * h)
        assert abs(fd - da[k - s1]) < 1e-6, (k, fd, da[k - s1])
        n += 1
    assert n > 0


def test_autograd_function_water_octamer_subset(gpu):
    """(H2O)8 even-tempered s,p,d: torch.autograd through RHFEnergy, checked
    against FD on a seeded subset of coordinates and exponents."""
    import torch
    from librint_amd import autograd, mol
    m = mol.water_cluster(8)
    slots = autograd.exponent_slots(m)
    coords = torch.tensor(m.coords(), dtype=torch.float64, requires_grad=True)
    exps = torch.tensor(m._env[slots], dtype=torch.float64, requires_grad=True)
    _say("autograd (H2O)8 nao", m.nao(), "exponents", len(slots))
    E = autograd.RHFEnergy.apply(coords, exps, m, slots)
    E.backward()
    gc = coords.grad.numpy()
    ge = exps.grad.numpy()
    rng = np.random.default_rng(11)
    csel = [(int(a), int(d)) for a, d in zip(rng.integers(0, len(m._atm), 4), rng.integers(0, 3, 4))]
    esel = [int(i) for i in rng.choice(len(slots), 4, replace=False)]

    def energy(c, e):
        with torch.no_grad():
            return float(autograd.RHFEnergy.apply(torch.tensor(c), torch.tensor(e), m, slots))

    c0 = coords.detach().numpy().copy()
    e0 = exps.detach().numpy().copy()
    for a, d in csel:
        def fc(x):
            c = c0.copy()
            c[a, d] = x
            return energy(c, e0)
        fd = _fd5(fc, c0[a, d], 1e-4)
        _say("  coord", a, d, fd, gc[a, d])
        assert abs(fd - gc[a, d]) < 1e-6
    for i in esel:
        def fe(x):
            e = e0.copy()
            e[i] = x
            return energy(c0, e)
        fd = _fd5(fe, e0[i], _exp_step(e0[i]))
        _say("  exponent", i, fd, ge[i])
        assert abs(fd - ge[i]) < 1e-6
