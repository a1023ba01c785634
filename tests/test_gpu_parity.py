"""HIP path vs the oracle through the C ABI (run on an MI355X)."""
import sys

import numpy as np
import pytest

from conftest import ref_mol, ref_truth
from oracle import refcpu

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rand_sym(n, seed):
    rng = np.random.default_rng(seed)
    P = rng.normal(size=(n, n))
    return 0.5 * (P + P.T)


def test_int2e_tensor_h2o_file_geometry(gpu):
    from librint_amd import scf
    _say("int2e h2o/sto3g file geometry")
    m = ref_mol("h2o", "sto3g", 10)
    R = scf.int2e(m)
    assert np.abs(R - refcpu.int2e(m)).max() < 1e-12


def test_int2e_tensor_h2o_baseline_geometry(gpu):
    from librint_amd import mol, scf
    _say("int2e h2o/sto3g O(0,0,0) H(+-0.757,0.587,0)")
    m = mol.h2o_sto3g()
    R = scf.int2e(m)
    assert np.abs(R - refcpu.int2e(m)).max() < 1e-12
    for perm in [(1, 0, 2, 3), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.abs(R - R.transpose(perm)).max() == 0.0


def test_shell_blocks_alkane_sampled(gpu):
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("shell blocks C16H34/6-31G sampled")
    m = mol.c16h34_631g()
    ctx = FockContext(m)
    rng = np.random.default_rng(5)
    qs = [tuple(int(v) for v in rng.integers(0, m.nbas, 4)) for _ in range(300)]
    # include near-diagonal and same-shell quartets
    qs += [(s, s, s, s) for s in range(0, m.nbas, 7)] + [(3, 4, 3, 4), (10, 2, 2, 10), (0, 1, 2, 3)]
    got = ctx.shells(qs)
    worst = 0.0
    for q, blk in zip(qs, got):
        worst = max(worst, np.abs(blk - refcpu.int2e_shell(m, q)).max())
    assert worst < 1e-12


@pytest.mark.parametrize("builder", ["h2o_file", "h2o"])
def test_fock_vs_oracle_full_loop(gpu, builder):
    from librint_amd import mol, scf
    _say("fock vs oracle", builder)
    m = {"h2o_file": lambda: ref_mol("h2o", "sto3g", 10), "h2o": mol.h2o_sto3g}[builder]()
    n = m.nao()
    for seed in (1, 2):
        P = _rand_sym(n, seed)
        G = scf.fock(m, P, 0.0)
        Gr = refcpu.fock2e(m, P)
        assert np.abs(G - Gr).max() < 1e-11 * max(1.0, np.abs(Gr).max())


def test_fock_vs_oracle_fixture_c2h6(gpu):
    import os
    from conftest import GOLD
    from librint_amd import mol, scf
    _say("fock vs oracle fixture C2H6/6-31G")
    d = np.load(os.path.join(GOLD, "fock_c2h6_631g.npz"))
    m = mol.alkane_631g(2)
    G = scf.fock(m, d["P"], 0.0)
    assert np.abs(G - d["G"]).max() < 1e-11 * np.abs(d["G"]).max()


def test_fock_c16h34_block_density(gpu):
    """Full-size check: a density supported on two shells; the oracle visits
    only the quartets that can touch it (reference loop restricted)."""
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("fock C16H34 block density")
    m = mol.c16h34_631g()
    ctx = FockContext(m, schwarz_tol=0.0)
    n = m.nao()
    bas = m._bas
    ao = np.concatenate([[0], np.cumsum([(b[1] + 1) * (b[1] + 2) // 2 for b in bas])])
    S = [40, 43]  # two shells on neighbouring carbons (s and p)
    P = np.zeros((n, n))
    rng = np.random.default_rng(9)
    idx = np.concatenate([np.arange(ao[s], ao[s + 1]) for s in S])
    B = rng.normal(size=(len(idx), len(idx)))
    P[np.ix_(idx, idx)] = B + B.T
    G = ctx.fock(P)
    nb = m.nbas
    qs = set()
    for i in range(nb):
        for j in range(nb):
            for a in S:
                for b in S:
                    qs.add((i, j, a, b))  # J: P[l,k] with k,l in S
                    qs.add((i, a, b, j))  # K: P[j,k] with j,k in S
    qs = np.array(sorted(qs), dtype=np.int32)
    Gr = refcpu.fock2e_sample(m, P, qs)
    assert np.abs(G - Gr).max() < 1e-11 * np.abs(Gr).max()
    assert np.abs(G - G.T).max() < 1e-13 * np.abs(G).max()


def test_fock_vs_screened_oracle_same_quartets(gpu):
    """At Schwarz 1e-12 the GPU build and the oracle's screened symmetric
    build visit the same canonical quartets, so G agrees to rounding."""
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("fock C4H10 vs screened oracle, tau 1e-12")
    m = mol.alkane_631g(4)
    ctx = FockContext(m, schwarz_tol=1e-12)
    P = _rand_sym(m.nao(), 11)
    G = ctx.fock(P)
    Gr, cnt = refcpu.fock2e_screened(m, P, 1e-12)
    assert cnt == ctx.stats()["quartets"]
    assert np.abs(G - Gr).max() < 1e-13 * np.abs(Gr).max()


def test_fock_shards_sum_to_whole_build(gpu):
    """One device's share of the build (the deal a multi-device run would
    make): the shares' G's sum to the whole G, their quartets to the whole
    count, and their costs are balanced."""
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("fock C16H34 shares of 4")
    m = mol.c16h34_631g()
    ctx = FockContext(m, schwarz_tol=1e-12)
    P = _rand_sym(m.nao(), 12)
    G = ctx.fock(P)
    whole = ctx.stats()
    parts, st = [], []
    for k in range(4):
        ctx.set_shard(4, k)
        parts.append(ctx.fock(P))
        st.append(ctx.stats())
    assert sum(s["quartets"] for s in st) == whole["quartets"]
    assert sum(s["prim_quartets"] for s in st) == whole["prim_quartets"]
    assert abs(sum(s["flops"] for s in st) - whole["flops"]) < 1e-9 * whole["flops"]
    assert np.abs(sum(parts) - G).max() < 1e-13 * np.abs(G).max()
    assert all(np.abs(g).max() > 0 for g in parts)
    fl = np.array([s["flops"] for s in st])
    assert fl.max() / fl.mean() < 1.05
    with pytest.raises(RuntimeError):
        ctx.set_shard(4, 4)
    ctx.set_shard(1, 0)
    assert np.abs(ctx.fock(P) - G).max() < 1e-13 * np.abs(G).max()


def test_fock_linearity_and_screening(gpu):
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("fock linearity C16H34")
    m = mol.c16h34_631g()
    ctx = FockContext(m, schwarz_tol=1e-12)
    n = m.nao()
    P1, P2 = _rand_sym(n, 3), _rand_sym(n, 4)
    G1, G2, G12 = ctx.fock(P1), ctx.fock(P2), ctx.fock(2.0 * P1 - 0.5 * P2)
    assert np.abs(G12 - (2.0 * G1 - 0.5 * G2)).max() < 1e-11 * np.abs(G12).max()
    st = ctx.stats()
    ctx.set_schwarz(0.0)
    G0 = ctx.fock(P1)
    assert ctx.stats()["quartets"] > st["quartets"]
    # Schwarz at 1e-12 on a density with |P| ~ 1: error bounded by tau * sum|P| per element
    assert np.abs(G0 - G1).max() < 1e-12 * np.abs(P1).sum()


@pytest.mark.parametrize("name,nelec", [("h2", 2), ("h2o", 10)])
def test_scf_c_truth_energy(gpu, name, nelec):
    from librint_amd import scf
    _say("scf_c truth", name)
    m = ref_mol(name, "sto3g", nelec)
    E = scf.scf(m, 200, 1e-6)
    assert abs(E - float(ref_truth(name, "sto3g_E"))) < 1e-10


def test_density_energy_and_baseline_geometry(gpu):
    from librint_amd import mol, scf
    _say("density/energy h2o baseline geometry")
    m = mol.h2o_sto3g()
    P = scf.density(m, 200, 1e-10)
    rc, Pr, it = refcpu.density(m, 200, 1e-10)
    assert rc == 0
    assert np.abs(P - Pr).max() < 1e-8
    E = scf.energy(m, P)
    assert abs(E - refcpu.energy(m, Pr)) < 1e-10
    assert abs(E - scf.scf(m, 200, 1e-10)) < 1e-10


def test_failures_are_reported(gpu):
    from librint_amd import scf
    _say("failure paths")
    m = ref_mol("h2o", "sto3g", 10)
    with pytest.raises(RuntimeError):
        scf.density(m, 1, 1e-12)  # cannot converge in one cycle
    with pytest.raises(RuntimeError):
        scf.scf(m, 1, 1e-12)
    from librint_amd import mol
    atoms = [(8, (0.0, 0.0, 0.0)), (1, (0.0, 1.4, 1.1)), (1, (0.0, -1.4, 1.1))]
    h_shell = mol.build(atoms, {8: mol.STO3G[8] + [(5, [0.8], [1.0])], 1: mol.STO3G[1]}, "h2o_h")
    with pytest.raises(RuntimeError):
        scf.int2e(h_shell)  # l = 5 is outside the GPU kernels


def test_fock_device_buffers(gpu):
    import torch
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("fock device buffers")
    m = mol.alkane_631g(4)
    ctx = FockContext(m, schwarz_tol=1e-14)
    P = _rand_sym(m.nao(), 8)
    Pd = torch.from_numpy(P).to(gpu)
    Gd = torch.empty_like(Pd)
    ctx.fock_device(Pd, Gd)
    torch.cuda.synchronize()
    G = ctx.fock(P)
    # atomics make the summation order vary run to run: equal to rounding only
    assert np.abs(Gd.cpu().numpy() - G).max() < 1e-13 * np.abs(G).max()


def test_d_shell_blocks_sampled(gpu):
    """Every angular class with d shells, including the cooperative (team)
    kernels, against the oracle on sampled shell quartets."""
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("d-shell blocks (H2O)2 even-tempered")
    m = mol.water_cluster(2)
    ctx = FockContext(m)
    L = m._bas[:, 1]
    by_l = {l: [i for i in range(m.nbas) if L[i] == l] for l in (0, 1, 2)}
    rng = np.random.default_rng(12)
    qs = []
    for la in (0, 1, 2):
        for lb in (0, 1, 2):
            for lc in (0, 1, 2):
                for ld in (0, 1, 2):
                    for _ in range(3):
                        qs.append(tuple(int(rng.choice(by_l[x])) for x in (la, lb, lc, ld)))
    d = by_l[2]
    qs += [(d[0], d[0], d[0], d[0]), (d[0], d[1], d[1], d[0]), (d[1], d[1], d[0], d[0])]
    got = ctx.shells(qs)
    for q, blk in zip(qs, got):
        ref = refcpu.int2e_shell(m, q)
        assert np.abs(blk - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), q


@pytest.mark.parametrize("builder", ["h2o_d", "h2o_def2svp"])
def test_fock_d_shells_vs_oracle(gpu, builder):
    from librint_amd import mol, scf
    _say("fock d shells", builder)
    m = {"h2o_d": mol.h2o_631gstar_like, "h2o_def2svp": lambda: ref_mol("h2o", "def2svp", 10)}[builder]()
    P = _rand_sym(m.nao(), 21)
    G = scf.fock(m, P, 0.0)
    Gr = refcpu.fock2e(m, P)
    assert np.abs(G - Gr).max() < 1e-11 * max(1.0, np.abs(Gr).max())


def test_tensor_d_shells(gpu):
    from librint_amd import mol, scf
    _say("int2e tensor with d shells")
    m = mol.h2o_631gstar_like()
    R = scf.int2e(m)
    Rr = refcpu.int2e(m)
    assert np.abs(R - Rr).max() < 1e-12
    # every image of a quartet carries the same value (one writer per image)
    for perm in [(1, 0, 2, 3), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.abs(R - R.transpose(perm)).max() == 0.0


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


def test_mixed_precision_fock(gpu):
    """FP32 recurrences below tau_mix: error bounded, energy error << 1 uEh."""
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("mixed precision C16H34")
    m = mol.c16h34_631g()
    ctx = FockContext(m, schwarz_tol=1e-12)
    P = _core_density(m)
    G64 = ctx.fock(P)
    ctx.set_precision(1, 0.0)          # tau_mix 0: every task stays FP64
    assert np.abs(ctx.fock(P) - G64).max() < 1e-12
    for tau in (1e-8, 1e-6, 1e-5):
        ctx.set_precision(1, tau)
        Gm = ctx.fock(P)
        fc = ctx.flag_counts()
        assert fc["quartets"][1] > 0
        dE = 0.5 * np.sum(P * (Gm - G64))
        _say("tau_mix", tau, "fp32 quartets", fc["quartets"], "dE", dE, "max dG", np.abs(Gm - G64).max())
        assert abs(dE) < 1e-6
        assert np.abs(Gm - G64).max() < 1e-5
    ctx.set_precision(0, 0.0, 1e-11)   # density screening only
    Gs = ctx.fock(P)
    assert ctx.flag_counts()["quartets"][2] > 0
    assert abs(0.5 * np.sum(P * (Gs - G64))) < 1e-7


@pytest.mark.parametrize("builder", ["c16h34", "w2"])
def test_stream_counts_agree(gpu, builder, monkeypatch):
    """The class launches after the first are dealt over several streams;
    any stream count gives the same G up to the summation order of the
    atomics (FP64 and mixed lists, slab and team classes).  Without
    LIBRINT_STREAMS a build of more than 12 class launches takes three
    streams, a smaller one two."""
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("stream counts", builder)
    m = {"c16h34": mol.c16h34_631g, "w2": lambda: mol.water_cluster(2)}[builder]()
    P = _core_density(m)
    out = {}
    for ns in ("1", "4", "default"):
        if ns == "default":
            monkeypatch.delenv("LIBRINT_STREAMS", raising=False)
        else:
            monkeypatch.setenv("LIBRINT_STREAMS", ns)
        ctx = FockContext(m, schwarz_tol=1e-12)
        out[ns] = ctx.fock(P)
        st = ctx.stats()
        want = min(int(ns), st["launches"] - 1) if ns != "default" else (3 if st["launches"] > 12 else 2)
        assert st["streams"] == max(1, want), (ns, st["streams"], st["launches"])
        ctx.set_precision(1, 1e-5)
        out[ns + "m"] = ctx.fock(P)
        ctx.close()
    scale = np.abs(out["1"]).max()
    for ns in ("4", "default"):
        assert np.abs(out["1"] - out[ns]).max() < 1e-13 * scale
        assert np.abs(out["1m"] - out[ns + "m"]).max() < 1e-13 * scale
    assert np.abs(out["1m"] - out["1"]).max() < 1e-6


def test_density_rows_in_lds_vs_oracle(gpu, monkeypatch):
    """Uncontracted water-cluster lists stage the bra's density rows in LDS
    (few primitive quartets per quartet); with and without the staging the
    build equals the oracle's screened build on the same quartets."""
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("density rows in LDS, (H2O)4 even-tempered")
    m = mol.water_cluster(4)
    P = _rand_sym(m.nao(), 13)
    Gr, cnt = refcpu.fock2e_screened(m, P, 1e-12)
    out = {}
    for mode in ("2", "0"):
        monkeypatch.setenv("LIBRINT_DROW", mode)
        ctx = FockContext(m, schwarz_tol=1e-12)
        out[mode] = ctx.fock(P)
        assert ctx.stats()["quartets"] == cnt
        staged = [c for c in ctx.class_stats() if c["drow"]]
        assert (len(staged) > 0) == (mode == "2")
        ctx.close()
    scale = np.abs(Gr).max()
    assert np.abs(out["2"] - Gr).max() < 1e-13 * scale
    assert np.abs(out["0"] - Gr).max() < 1e-13 * scale


def test_slab_budget_chunks_vs_oracle(gpu, monkeypatch):
    """A 1 MiB Coulomb-slab budget (LIBRINT_SLAB_MB) splits every slab class
    launch into many chunks, each reduced after it: the build still equals
    the oracle's screened build on the same quartets, in FP64 and with the
    mixed kernels."""
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("slab budget 1 MiB, (H2O)4 even-tempered")
    m = mol.water_cluster(4)
    P = _rand_sym(m.nao(), 17)
    Gr, cnt = refcpu.fock2e_screened(m, P, 1e-12)
    scale = np.abs(Gr).max()
    out = {}
    for mb in ("1", "1024"):
        monkeypatch.setenv("LIBRINT_SLAB_MB", mb)
        ctx = FockContext(m, schwarz_tol=1e-12)
        out[mb] = ctx.fock(P)
        assert ctx.stats()["quartets"] == cnt
        ctx.set_precision(1, 1e-7)
        out[mb + "m"] = ctx.fock(P)
        ctx.close()
    assert np.abs(out["1"] - Gr).max() < 1e-13 * scale
    assert np.abs(out["1024"] - Gr).max() < 1e-13 * scale
    assert np.abs(out["1m"] - out["1024m"]).max() < 1e-13 * scale
    assert np.abs(out["1m"] - Gr).max() < 1e-6


def test_chunked_classes_vs_team_and_oracle(gpu, monkeypatch):
    """The d-shell classes whose contracted block exceeds a lane's registers
    ((12|11), (12|12), (22|11), (22|02)) run lane per quartet with the block
    evaluated in chunks inside the group kernel; without the group kernel
    (LIBRINT_NO_GROUP) they fall back to the team kernel.  Both builds, and
    the chunked one with mixed precision, equal the oracle's screened build."""
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("chunked d classes vs team fallback, (H2O)4 even-tempered")
    m = mol.water_cluster(4)
    P = _rand_sym(m.nao(), 23)
    Gr, cnt = refcpu.fock2e_screened(m, P, 1e-12)
    scale = np.abs(Gr).max()
    chunked = {(1, 2, 1, 1), (1, 2, 1, 2), (2, 2, 1, 1), (2, 2, 0, 2)}
    out = {}
    for mode in ("group", "flat"):
        if mode == "flat":
            monkeypatch.setenv("LIBRINT_NO_GROUP", "1")
        ctx = FockContext(m, schwarz_tol=1e-12)
        out[mode] = ctx.fock(P)
        assert ctx.stats()["quartets"] == cnt
        kern = {tuple(int(c[k]) for k in ("la", "lb", "lc", "ld")): c["kernel"] for c in ctx.class_stats()}
        assert chunked <= set(kern)
        for cl in chunked:
            assert kern[cl].startswith("jk_group_kernel" if mode == "group" else "jk_team_kernel"), kern[cl]
        if mode == "group":
            ctx.set_precision(1, 1e-6)
            out["mixed"] = ctx.fock(P)
        ctx.close()
    assert np.abs(out["group"] - Gr).max() < 1e-13 * scale
    assert np.abs(out["flat"] - Gr).max() < 1e-13 * scale
    assert np.abs(out["mixed"] - Gr).max() < 1e-6


def test_mixed_precision_d_shells(gpu):
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("mixed precision (H2O)4 even-tempered")
    m = mol.water_cluster(4)
    ctx = FockContext(m, schwarz_tol=1e-12)
    P = _core_density(m)
    G64 = ctx.fock(P)
    ctx.set_precision(1, 1e-6)
    Gm = ctx.fock(P)
    assert ctx.flag_counts()["quartets"][1] > 0
    assert abs(0.5 * np.sum(P * (Gm - G64))) < 1e-6


@pytest.mark.parametrize("builder", ["w4", "c16h34"])
def test_scf_purification_matches_eigh(gpu, builder, monkeypatch):
    """In the DIIS mode the SCF takes its densities from the purified
    occupied-space projector instead of diagonalising F until the density
    changes by less than 100 conv; it converges to the same energy (1e-10 Eh)
    and density in the same number of cycles as the eigensolver path
    (LIBRINT_SCF_EIGH=1)."""
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("SCF purification vs eigh", builder)
    m = {"w4": lambda: mol.water_cluster(4), "c16h34": mol.c16h34_631g}[builder]()
    out = {}
    for mode in ("purify", "eigh"):
        if mode == "eigh":
            monkeypatch.setenv("LIBRINT_SCF_EIGH", "1")
        ctx = FockContext(m, schwarz_tol=1e-12)
        E, P, it = ctx.scf(imax=100, conv=1e-6, guess=1, diis=True, conv_e=1e-8)
        out[mode] = (E, P, it, ctx.stats()["scf_purified"])
        ctx.close()
    (E1, P1, it1, n1), (E2, P2, it2, n2) = out["purify"], out["eigh"]
    _say("E", E1, E2, "iterations", it1, it2, "purified", n1, n2)
    assert n2 == 0 and 3 <= n1 <= it1  # the guess and the early cycles purified
    assert abs(E1 - E2) < 1e-10
    assert abs(it1 - it2) <= 1
    assert np.abs(P1 - P2).max() < 1e-6


def test_mixed_precision_energy_vs_oracle(gpu):
    """Mixed-precision two-electron energies against the oracle's CPU build
    (Schwarz-screened restatement of integral2e_fock, FP64) on the (H2O)4
    even-tempered cluster with its core-Hamiltonian density: at every
    tau_mix of the bench range the energy error stays within 1 uEh, and the
    FP64 GPU build's error is at rounding level."""
    from librint_amd import mol
    from librint_amd.context import FockContext
    _say("mixed precision energy vs oracle, (H2O)4 even-tempered")
    m = mol.water_cluster(4)
    P = _core_density(m)
    Gr, _ = refcpu.fock2e_screened(m, P, 1e-12)
    E_ref = 0.5 * np.sum(P * Gr)
    ctx = FockContext(m, schwarz_tol=1e-12)
    E64 = 0.5 * np.sum(P * ctx.fock(P))
    assert abs(E64 - E_ref) < 1e-10 * abs(E_ref)
    for tau in (1e-8, 1e-7, 1e-6, 1e-5):
        ctx.set_precision(1, tau)
        Em = 0.5 * np.sum(P * ctx.fock(P))
        _say("tau_mix", tau, "fp32 quartets", ctx.flag_counts()["quartets"][1], "E - E_oracle", Em - E_ref)
        assert abs(Em - E_ref) < 1e-6
    ctx.close()


def _c2s_independent():
    """cart -> real spherical rows for l = 2 fitted from scipy's spherical
    harmonics (r^2 Y_2m on random points, least squares over the six
    monomials), sign of each row fixed by its largest coefficient."""
    import scipy.special as sp
    rng = np.random.default_rng(4)
    v = rng.normal(size=(400, 3))
    x, y, z = v.T
    r = np.linalg.norm(v, axis=1)
    th, ph = np.arccos(z / r), np.arctan2(y, x)
    Y = lambda m: (sp.sph_harm_y(2, m, th, ph) if hasattr(sp, "sph_harm_y") else sp.sph_harm(m, 2, ph, th))
    real = {0: Y(0).real}
    for m in (1, 2):
        real[m] = np.sqrt(2) * (-1) ** m * Y(m).real
        real[-m] = np.sqrt(2) * (-1) ** m * Y(m).imag
    mono = np.stack([x * x, x * y, x * z, y * y, y * z, z * z], axis=1)
    rows = []
    for m in (-2, -1, 0, 1, 2):
        c, *_ = np.linalg.lstsq(mono, real[m] * r * r, rcond=None)
        c[np.abs(c) < 1e-12] = 0.0
        rows.append(c * np.sign(c[np.argmax(np.abs(c))]))
    return np.array(rows)


def test_spherical_output(gpu):
    """coord = 1: identical to cartesian for s,p bases; identity on-site
    overlap for a normalised d shell; int2e equal to the cartesian tensor
    transformed with independently fitted solid harmonics."""
    from librint_amd import mol, scf
    m = ref_mol("h2o", "sto3g", 10)
    for typ in ("ovlp", "kin", "nuc"):
        np.testing.assert_array_equal(scf.int1e(m, typ, "sph"), scf.int1e(m, typ, "cart"))
    np.testing.assert_array_equal(scf.int2e(m, "sph"), scf.int2e(m, "cart"))
    md = mol.h2o_631gstar_like()
    Ssph = scf.int1e(md, "ovlp", "sph")
    # O shells: 1s, 2s, 2p, d -> the d block is spherical functions 5..9
    np.testing.assert_allclose(Ssph[5:10, 5:10], np.eye(5), atol=1e-12)
    # transform matrix: identity for s/p, fitted rows for the d shell
    ncart, nsph = md.nao(0), md.nao(1)
    C = np.zeros((nsph, ncart))
    C[:5, :5] = np.eye(5)
    C[5:10, 5:11] = _c2s_independent()
    C[10:, 11:] = np.eye(nsph - 10)
    for typ in ("ovlp", "kin", "nuc"):
        np.testing.assert_allclose(scf.int1e(md, typ, "sph"), C @ scf.int1e(md, typ, "cart") @ C.T, atol=1e-12)
    Rc = scf.int2e(md, "cart")
    Rs = scf.int2e(md, "sph")
    ref = np.einsum("ai,bj,ck,dl,ijkl->abcd", C, C, C, C, Rc, optimize=True)
    assert Rs.shape == (nsph,) * 4
    assert np.abs(Rs - ref).max() < 1e-12
