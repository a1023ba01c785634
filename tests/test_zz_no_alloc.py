"""librint_fock_device_noalloc (the Fock build on the library's static
device scratch, no allocation) against librint_fock_device through the C
ABI: two calls in a row on different contexts, the second with a smaller
basis, each matching the allocating entry point; mixed precision and
density screening through the same scratch; and the size limit."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(ctx, P):
    import torch
    from librint_amd._bindings import library
    Pd = torch.from_numpy(np.ascontiguousarray(P)).to("cuda:0")
    G1 = torch.full_like(Pd, 7.0)
    G2 = torch.full_like(Pd, -3.0)
    vp = ctypes.c_void_p
    assert library.librint_fock_device(ctx.h, vp(Pd.data_ptr()), vp(G1.data_ptr()), None) == 0
    torch.cuda.synchronize()
    rc = library.librint_fock_device_noalloc(ctx.h, vp(Pd.data_ptr()), vp(G2.data_ptr()), None)
    torch.cuda.synchronize()
    assert rc == 0, library.librint_last_error()
    return G1.cpu().numpy(), G2.cpu().numpy()


def _rand_sym(n, seed):
    rng = np.random.default_rng(seed)
    P = rng.normal(size=(n, n))
    return 0.5 * (P + P.T)


def test_noalloc_matches_allocating_entry(gpu):
    from librint_amd import mol
    from librint_amd.context import FockContext
    from oracle import refcpu
    cases = [("C6H14/6-31G", mol.alkane_631g(6), 1e-12), ("(H2O)2 s,p,d", mol.water_cluster(2), 1e-12),
             ("H2O/STO-3G", mol.h2o_sto3g(), 0.0), ("H2O f,g", mol.h2o_fg(), 0.0)]
    for k, (name, m, tau) in enumerate(cases):
        ctx = FockContext(m, schwarz_tol=tau)
        P = _rand_sym(m.nao(), 40 + k)
        for rep in range(2):
            G1, G2 = _both(ctx, P)
            scale = np.abs(G1).max()
            _say(name, "call", rep, "nao", m.nao(), "max |noalloc - alloc|", np.abs(G2 - G1).max(), "scale", scale)
            assert np.abs(G2 - G1).max() < 1e-12 * scale
        if m.nao() <= 60:
            Gr = refcpu.fock2e(m, P)
            assert np.abs(G2 - Gr).max() < 1e-11 * np.abs(Gr).max()
        if name.startswith("C6"):
            # density screening and mixed precision use more of the scratch
            ctx.set_precision(1, 1e-6, 1e-11)
            G1, G2 = _both(ctx, P)
            assert np.abs(G2 - G1).max() < 1e-12 * np.abs(G1).max()
            ctx.set_precision(0, 0.0, 0.0)
        ctx.close()


def test_noalloc_rejects_builds_beyond_its_scratch(gpu):
    """3000 s functions on far-apart centres: nao^2 doubles (72 MB) exceed
    the 64 MiB buffer, so the entry refuses with a message."""
    import torch
    from librint_amd import mol
    from librint_amd._bindings import library
    from librint_amd.context import FockContext
    atoms = [(1, (60.0 * i, 0.0, 0.0)) for i in range(3000)]
    m = mol.build(atoms, {1: [(0, [1.0], [1.0])]}, "h3000", normalized=False)
    ctx = FockContext(m, schwarz_tol=1e-12)
    n = m.nao()
    Pd = torch.zeros((n, n), dtype=torch.float64, device="cuda:0")
    Gd = torch.zeros_like(Pd)
    vp = ctypes.c_void_p
    assert library.librint_fock_device_noalloc(ctx.h, vp(Pd.data_ptr()), vp(Gd.data_ptr()), None) == -1
    assert b"scratch" in library.librint_last_error()
    ctx.close()
