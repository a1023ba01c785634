"""Learned SCF starting density: a small MLP maps atom-centred geometry
features of every shell pair to that pair's density-matrix block.

Block model (rotation covariant): for shells A on atom i and B on atom j,
with u the unit vector from i to j (for i == j: the direction of the
atom's environment vector v_i = sum_k f(r_ik)(R_k - R_i)/r_ik),
    s-s:  P_AB = t0
    s-p:  P_AB[x] = t0 u_x          (p-s likewise, along the same u)
    p-p:  P_AB = t0 u u^T + t1 (1 - u u^T)
The network predicts (t0, t1) from one-hot shell kinds, a radial basis of
r_ij, smooth C/H coordination counts of both atoms and the environment
projections; targets are the least-squares (t0, t1) of converged densities.

Training and inference run on the GPU through the library's MLP kernels
(MFMA BF16 GEMMs with FP32 accumulation, ReLU backward, column sums, masked
MSE, Adam on FP32 master weights); torch only provides device buffers.  The
predicted density is only a starting point: the SCF (core-Hamiltonian guess
as the fallback) converges to the same energy either way, so the network can
change the iteration count and nothing else.
"""
import ctypes
import math
import time

import numpy as np

from . import mol as M

ANG = M.BOHR

# shell kinds of the 6-31G basis: (element, index of the shell in its list)
KINDS = [(1, 0), (1, 1), (6, 0), (6, 1), (6, 2), (6, 3), (6, 4)]
KIND_INDEX = {k: i for i, k in enumerate(KINDS)}
NRAD = 13
NFEAT = 48   # padded input width (multiple of 16)
NOUT = 16    # padded output width; columns 0, 1 are t0, t1
HIDDEN = 256


# ----------------------------------------------------------- geometries --

def _rotate(points, origin, axis, angle):
    axis = axis / np.linalg.norm(axis)
    p = points - origin
    c, s = math.cos(angle), math.sin(angle)
    return origin + p * c + np.cross(axis, p) * s + np.outer(p @ axis, axis) * (1 - c)


def perturbed_alkane(n_carbon, rng, sigma=0.05, torsions=True, dmin=1.6):
    """CnH2n+2 (angstrom): all-trans start, uniform random torsion about every
    C-C bond, Gaussian displacement sigma on every atom; redrawn while two
    atoms that are not bonded come closer than dmin angstrom."""
    C0, H0 = M.alkane_geometry(n_carbon)
    owner = np.array([int(np.argmin(np.linalg.norm(C0 - h, axis=1))) for h in H0])
    for _ in range(100):
        C, H = C0.copy(), H0.copy()
        if torsions:
            for b in range(n_carbon - 1):
                ang = rng.uniform(-math.pi, math.pi)
                axis = C[b + 1] - C[b]
                sel_c = np.arange(n_carbon) > b + 1
                sel_h = owner > b
                C[sel_c] = _rotate(C[sel_c], C[b], axis, ang)
                H[sel_h] = _rotate(H[sel_h], C[b], axis, ang)
        C = C + rng.normal(scale=sigma, size=C.shape)
        H = H + rng.normal(scale=sigma, size=H.shape)
        X = np.vstack([C, H])
        d = np.linalg.norm(X[:, None] - X[None], axis=-1)
        np.fill_diagonal(d, 10.0)
        bonded = np.zeros_like(d, dtype=bool)  # C-C 1.54, C-H 1.09 angstrom
        bonded[:n_carbon, :n_carbon] = d[:n_carbon, :n_carbon] < 1.7
        bonded[:n_carbon, n_carbon:] = d[:n_carbon, n_carbon:] < 1.25
        bonded[n_carbon:, :n_carbon] = d[n_carbon:, :n_carbon] < 1.25
        if np.all(d[~bonded] >= dmin):
            break
    atoms = [(6, c) for c in C] + [(1, h) for h in H]
    return atoms


def alkane_sample(n_carbon, rng):
    atoms = perturbed_alkane(n_carbon, rng)
    return M.build([(Z, tuple(float(v) / ANG for v in x)) for Z, x in atoms], M.B631G,
                   "C%dH%d_pert" % (n_carbon, 2 * n_carbon + 2))


# ------------------------------------------------------- featurization --

class ShellInfo:
    """Per-shell metadata of a 6-31G molecule (segmented, one contraction)."""

    def __init__(self, m):
        bas = np.asarray(m._bas)
        atm = np.asarray(m._atm)
        self.coords = m.coords()
        self.Z = atm[:, 0]
        self.atom, self.l, self.kind, self.ao = [], [], [], []
        seen = {}
        ao = 0
        for b in bas:
            a, l, nctr = int(b[0]), int(b[1]), int(b[3])
            if nctr != 1 or l > 1:
                raise ValueError("the learned guess covers segmented s,p shells (6-31G)")
            k = seen.get(a, 0)
            seen[a] = k + 1
            key = (int(self.Z[a]), k)
            if key not in KIND_INDEX:
                raise ValueError("shell kind %r outside the 6-31G C/H set" % (key,))
            self.atom.append(a)
            self.l.append(l)
            self.kind.append(KIND_INDEX[key])
            self.ao.append(ao)
            ao += 1 if l == 0 else 3
        self.nao = ao
        self.atom = np.array(self.atom)
        self.l = np.array(self.l)
        self.kind = np.array(self.kind)
        self.ao = np.array(self.ao)


def _env(coords, Z):
    """Smooth C and H coordination counts and environment vectors per atom."""
    d = coords[:, None] - coords[None]
    r = np.linalg.norm(d, axis=-1)
    np.fill_diagonal(r, np.inf)
    f = 1.0 / (1.0 + np.exp(4.0 * (r - 3.4)))  # ~1 within bonding distance (bohr)
    nC = (f * (Z[None] == 6)).sum(1)
    nH = (f * (Z[None] == 1)).sum(1)
    w = np.exp(-(r / 3.0) ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        unit = np.where(np.isfinite(r)[..., None], -d / r[..., None], 0.0)
    v = (w[..., None] * unit).sum(1)
    return nC, nH, v


def _pairs(si):
    ns = len(si.l)
    A, B = np.triu_indices(ns)
    return A, B


def _directions(si, A, B, v):
    ia, ib = si.atom[A], si.atom[B]
    d = si.coords[ib] - si.coords[ia]
    r = np.linalg.norm(d, axis=1)
    onsite = ia == ib
    u = np.zeros_like(d)
    u[~onsite] = d[~onsite] / r[~onsite, None]
    vn = np.linalg.norm(v[ia], axis=1)
    ok = onsite & (vn > 1e-8)
    u[ok] = v[ia][ok] / vn[ok, None]
    return r, onsite, u


def features(m, si=None):
    """(X float32 [npair, NFEAT], (A, B, u)) for every shell pair A <= B."""
    si = si or ShellInfo(m)
    A, B = _pairs(si)
    nC, nH, v = _env(si.coords, si.Z)
    r, onsite, u = _directions(si, A, B, v)
    ia, ib = si.atom[A], si.atom[B]
    X = np.zeros((len(A), NFEAT), dtype=np.float32)
    X[np.arange(len(A)), si.kind[A]] = 1.0
    X[np.arange(len(A)), 7 + si.kind[B]] = 1.0
    X[:, 14] = onsite
    mu = np.arange(NRAD, dtype=np.float64)
    X[:, 15:15 + NRAD] = np.exp(-((r[:, None] - mu[None]) / 0.7) ** 2)
    c = 15 + NRAD
    X[:, c + 0] = nC[ia] / 4.0
    X[:, c + 1] = nH[ia] / 4.0
    X[:, c + 2] = nC[ib] / 4.0
    X[:, c + 3] = nH[ib] / 4.0
    X[:, c + 4] = np.linalg.norm(v[ia], axis=1)
    X[:, c + 5] = np.linalg.norm(v[ib], axis=1)
    X[:, c + 6] = np.where(onsite, 0.0, (v[ia] * u).sum(1))
    X[:, c + 7] = np.where(onsite, 0.0, -(v[ib] * u).sum(1))
    return X, (A, B, u)


def targets(m, P, si=None, geo=None):
    """Least-squares block coefficients (T [npair, 2], mask [npair, 2])."""
    si = si or ShellInfo(m)
    if geo is None:
        _, geo = features(m, si)
    A, B, u = geo
    T = np.zeros((len(A), 2), dtype=np.float32)
    W = np.zeros((len(A), 2), dtype=np.float32)
    for k in range(len(A)):
        a, b = A[k], B[k]
        la, lb = si.l[a], si.l[b]
        blk = P[si.ao[a]:si.ao[a] + (1 if la == 0 else 3), si.ao[b]:si.ao[b] + (1 if lb == 0 else 3)]
        if la == 0 and lb == 0:
            T[k, 0] = blk[0, 0]
            W[k, 0] = 1.0
        elif la == 0:
            T[k, 0] = blk[0] @ u[k]
            W[k, 0] = 1.0
        elif lb == 0:
            T[k, 0] = blk[:, 0] @ u[k]
            W[k, 0] = 1.0
        else:
            t0 = u[k] @ blk @ u[k]
            T[k, 0] = t0
            T[k, 1] = 0.5 * (np.trace(blk) - t0)
            W[k] = 1.0
    return T, W


def assemble(m, t, si=None, geo=None):
    """Density from block coefficients t [npair, >= 2]."""
    si = si or ShellInfo(m)
    if geo is None:
        _, geo = features(m, si)
    A, B, u = geo
    P = np.zeros((si.nao, si.nao))
    for k in range(len(A)):
        a, b = A[k], B[k]
        la, lb = si.l[a], si.l[b]
        t0, t1 = float(t[k, 0]), float(t[k, 1])
        if la == 0 and lb == 0:
            blk = np.array([[t0]])
        elif la == 0:
            blk = (t0 * u[k])[None, :]
        elif lb == 0:
            blk = (t0 * u[k])[:, None]
        else:
            uu = np.outer(u[k], u[k])
            blk = t0 * uu + t1 * (np.eye(3) - uu)
        ra = slice(si.ao[a], si.ao[a] + blk.shape[0])
        rb = slice(si.ao[b], si.ao[b] + blk.shape[1])
        P[ra, rb] = blk
        P[rb, ra] = blk.T
    return P


# ------------------------------------------------------------ network --

def _lib():
    from ._bindings import library
    return library


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _check(rc, what):
    if rc != 0:
        from .utils import last_error
        raise RuntimeError("%s failed: %s" % (what, last_error()))


class MLP:
    """NFEAT -> 3 x HIDDEN (ReLU) -> NOUT, BF16 weights and activations,
    FP32 master weights and accumulation, Adam.  Layer l: Z = X W_l + b_l."""

    def __init__(self, seed=0, device="cuda:0", dims=(NFEAT, HIDDEN, HIDDEN, HIDDEN, NOUT)):
        import torch
        init_device(device)
        self.torch = torch
        self.dev = torch.device(device)
        self.dims = dims
        g = torch.Generator().manual_seed(seed)
        self.W, self.Wb, self.b, self.mW, self.vW, self.mb, self.vb = [], [], [], [], [], [], []
        for i in range(len(dims) - 1):
            w = torch.randn(dims[i], dims[i + 1], generator=g) * math.sqrt(2.0 / dims[i])
            if i == len(dims) - 2:
                w *= 0.1
            self.W.append(w.to(self.dev).contiguous())
            self.Wb.append(torch.empty(dims[i], dims[i + 1], dtype=torch.bfloat16, device=self.dev))
            self.b.append(torch.zeros(dims[i + 1], device=self.dev))
            for lst in (self.mW, self.vW):
                lst.append(torch.zeros_like(self.W[-1]))
            for lst in (self.mb, self.vb):
                lst.append(torch.zeros_like(self.b[-1]))
        self.step_count = 0
        self.refresh_bf16()

    def stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.dev).cuda_stream)

    def refresh_bf16(self):
        for w, wb in zip(self.W, self.Wb):
            _check(_lib().librint_mlp_to_bf16(w.numel(), _ptr(w), _ptr(wb), self.stream()), "to_bf16")

    def to_bf16(self, x):
        y = self.torch.empty(x.shape, dtype=self.torch.bfloat16, device=self.dev)
        _check(_lib().librint_mlp_to_bf16(x.numel(), _ptr(x), _ptr(y), self.stream()), "to_bf16")
        return y

    def forward(self, xb, keep=False):
        """xb: BF16 [rows, NFEAT] on the device.  Returns FP32 [rows, NOUT]."""
        torch = self.torch
        rows = xb.shape[0]
        acts = [(None, xb)]
        h = xb
        nl = len(self.W)
        for i in range(nl):
            n_out = self.dims[i + 1]
            last = i == nl - 1
            z = torch.empty(rows, n_out, device=self.dev)
            zb = None if last else torch.empty(rows, n_out, dtype=torch.bfloat16, device=self.dev)
            _check(_lib().librint_mlp_gemm(rows, n_out, self.dims[i], _ptr(h), self.dims[i], 0, _ptr(self.Wb[i]), n_out,
                                           0, _ptr(z), n_out, _ptr(self.b[i]), 0 if last else 1, 0, _ptr(zb), n_out,
                                           self.stream()), "gemm")
            acts.append((z, zb))
            h = zb
        if keep:
            self._acts = acts
        return acts[-1][0]

    def backward(self, dy_f, dy_b):
        """Gradients (dW, db) for the output gradient dy (FP32 and BF16)."""
        torch = self.torch
        acts = self._acts
        rows = dy_f.shape[0]
        nl = len(self.W)
        dW, db = [None] * nl, [None] * nl
        dpre_f, dpre_b = dy_f, dy_b
        for i in reversed(range(nl)):
            n_in, n_out = self.dims[i], self.dims[i + 1]
            x_b = acts[i][1]
            dW[i] = torch.empty(n_in, n_out, device=self.dev)
            _check(_lib().librint_mlp_gemm(n_in, n_out, rows, _ptr(x_b), n_in, 1, _ptr(dpre_b), n_out, 0, _ptr(dW[i]),
                                           n_out, None, 0, 0, None, 0, self.stream()), "gemm dW")
            db[i] = torch.empty(n_out, device=self.dev)
            _check(_lib().librint_mlp_colsum(rows, n_out, _ptr(dpre_f), n_out, _ptr(db[i]), 0, self.stream()), "colsum")
            if i == 0:
                break
            dz = torch.empty(rows, n_in, device=self.dev)
            _check(_lib().librint_mlp_gemm(rows, n_in, n_out, _ptr(dpre_b), n_out, 0, _ptr(self.Wb[i]), n_out, 1,
                                           _ptr(dz), n_in, None, 0, 0, None, 0, self.stream()), "gemm dX")
            h_f = acts[i][0]
            dpre_f = torch.empty(rows, n_in, device=self.dev)
            dpre_b = torch.empty(rows, n_in, dtype=torch.bfloat16, device=self.dev)
            _check(_lib().librint_mlp_relu_bwd(dz.numel(), _ptr(dz), _ptr(h_f), _ptr(dpre_f), _ptr(dpre_b),
                                               self.stream()), "relu_bwd")
        return dW, db

    def loss_grad(self, y, t, mask, scale):
        torch = self.torch
        rows = y.shape[0]
        dy_f = torch.empty_like(y)
        dy_b = torch.empty(y.shape, dtype=torch.bfloat16, device=self.dev)
        loss = torch.zeros(1, dtype=torch.float64, device=self.dev)
        _check(_lib().librint_mlp_mse(rows, t.shape[1], _ptr(y), y.shape[1], _ptr(t), _ptr(mask), t.shape[1],
                                      float(scale), _ptr(dy_f), _ptr(dy_b), _ptr(loss), self.stream()), "mse")
        return loss, dy_f, dy_b

    def adam(self, dW, db, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        L = _lib()
        for i in range(len(self.W)):
            _check(L.librint_mlp_adam(self.W[i].numel(), _ptr(self.W[i]), _ptr(dW[i]), _ptr(self.mW[i]),
                                      _ptr(self.vW[i]), lr, b1, b2, eps, self.step_count, _ptr(self.Wb[i]),
                                      self.stream()), "adam")
            _check(L.librint_mlp_adam(self.b[i].numel(), _ptr(self.b[i]), _ptr(db[i]), _ptr(self.mb[i]),
                                      _ptr(self.vb[i]), lr, b1, b2, eps, self.step_count, None, self.stream()), "adam")

    def train_step(self, xb, t, mask, lr):
        y = self.forward(xb, keep=True)
        scale = max(float(mask.sum().item()), 1.0) if isinstance(mask, self.torch.Tensor) else float(mask)
        loss, dy_f, dy_b = self.loss_grad(y, t, mask, scale)
        dW, db = self.backward(dy_f, dy_b)
        self.adam(dW, db, lr=lr)
        return loss

    def predict(self, X):
        """numpy float32 [rows, NFEAT] -> numpy [rows, 2]."""
        torch = self.torch
        x = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).to(self.dev)
        y = self.forward(self.to_bf16(x))
        return y[:, :2].cpu().numpy()


# ------------------------------------------------------ data and driver --

def init_device(device="cuda:0"):
    """Initialise torch's device before the library touches the GPU: torch
    carries its own HIP runtime, which fails to start once the library's
    runtime has initialised the device first."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the learned guess needs a GPU")
    torch.zeros(1, device=device)


def converged_density(m, conv=1e-6, conv_e=1e-10, schwarz=1e-12):
    from .context import FockContext
    ctx = FockContext(m, schwarz)
    E, P, it = ctx.scf(imax=200, conv=conv, guess=1, diis=True, conv_e=conv_e)
    ctx.close()
    if not np.isfinite(E):
        raise RuntimeError("SCF did not converge for " + m.name)
    return E, P, it


def make_dataset(n_mol, seed, carbons=(2, 3, 4, 5, 6), log=None):
    """Seeded perturbed alkanes with their converged GPU-SCF densities."""
    init_device()
    rng = np.random.default_rng(seed)
    X, T, W, mols = [], [], [], []
    t0 = time.time()
    for i in range(n_mol):
        nc = int(rng.choice(carbons))
        m = alkane_sample(nc, rng)
        E, P, it = converged_density(m)
        si = ShellInfo(m)
        x, geo = features(m, si)
        t, w = targets(m, P, si, geo)
        X.append(x)
        T.append(t)
        W.append(w)
        mols.append(m)
        if log and (i + 1) % 256 == 0:
            log("  %d/%d molecules (%.1f s)" % (i + 1, n_mol, time.time() - t0))
    return np.vstack(X), np.vstack(T), np.vstack(W), mols


def train(X, T, Wm, epochs=40, batch=8192, lr=1e-3, seed=0, log=None):
    import torch
    net = MLP(seed=seed)
    dev = net.dev
    x = torch.from_numpy(np.ascontiguousarray(X)).to(dev)
    xb = net.to_bf16(x)
    del x
    t = torch.from_numpy(np.ascontiguousarray(T)).to(dev)
    w = torch.from_numpy(np.ascontiguousarray(Wm)).to(dev)
    n = X.shape[0]
    g = torch.Generator(device=dev).manual_seed(seed)
    history = []
    steps = 0
    total = epochs * ((n + batch - 1) // batch)
    for ep in range(epochs):
        perm = torch.randperm(n, device=dev, generator=g)
        tot = 0.0
        for s0 in range(0, n, batch):
            idx = perm[s0:s0 + batch]
            # cosine learning-rate decay
            lr_s = lr * 0.5 * (1.0 + math.cos(math.pi * steps / total))
            loss = net.train_step(xb[idx].contiguous(), t[idx].contiguous(), w[idx].contiguous(), lr_s)
            tot += float(loss.item()) * len(idx)
            steps += 1
        history.append(tot / n)
        if log and (ep % 5 == 4 or ep == epochs - 1):
            log("  epoch %d mse %.3e" % (ep + 1, history[-1]))
    return net, history


def guess_density(net, m):
    si = ShellInfo(m)
    X, geo = features(m, si)
    t = net.predict(X)
    return assemble(m, t, si, geo)


def scf_iterations(m, P0=None, conv=1e-5, conv_e=1e-8, schwarz=1e-12):
    """(energy, iterations) from the core-Hamiltonian guess (P0 None) or P0."""
    from .context import FockContext
    ctx = FockContext(m, schwarz)
    if P0 is None:
        E, P, it = ctx.scf(imax=200, conv=conv, guess=1, diis=True, conv_e=conv_e)
    else:
        E, P, it = ctx.scf(imax=200, conv=conv, guess=2, diis=True, P0=P0, conv_e=conv_e)
    ctx.close()
    return E, it


def evaluate(net, mols, log=None):
    rows = []
    for m in mols:
        E0, i0 = scf_iterations(m)
        E1, i1 = scf_iterations(m, guess_density(net, m))
        rows.append((m.name, E0, i0, E1, i1))
    it0 = np.array([r[2] for r in rows], dtype=float)
    it1 = np.array([r[4] for r in rows], dtype=float)
    dE = np.array([abs(r[3] - r[1]) for r in rows])
    return {"molecules": len(rows), "iterations_core_h": float(it0.mean()), "iterations_mlp": float(it1.mean()),
            "reduction": float(1.0 - it1.sum() / it0.sum()), "max_energy_difference_Eh": float(dE.max()),
            "converged_all": bool(np.all(np.isfinite([r[1] for r in rows] + [r[3] for r in rows])))}


def run(n_train, n_test, epochs=40, seed=2024, log=None):
    """The config-3 experiment: train on n_train perturbed C2-C6 alkanes
    (converged GPU-SCF densities), then compare SCF iteration counts from the
    network's guess and from the core-Hamiltonian guess on n_test held-out
    molecules drawn from the seed + 1 stream."""
    t0 = time.time()
    if log:
        log("training set: %d molecules" % n_train)
    X, T, W, _ = make_dataset(n_train, seed, log=log)
    t_data = time.time() - t0
    if log:
        log("pairs %d; training" % X.shape[0])
    t1 = time.time()
    net, hist = train(X, T, W, epochs=epochs, log=log)
    t_train = time.time() - t1
    if log:
        log("held-out set: %d molecules" % n_test)
    rng = np.random.default_rng(seed + 1)
    tests = [alkane_sample(int(rng.choice((2, 3, 4, 5, 6))), rng) for _ in range(n_test)]
    res = evaluate(net, tests, log=log)
    res.update({"train_molecules": n_train, "train_pairs": int(X.shape[0]), "epochs": epochs,
                "first_train_mse": hist[0], "final_train_mse": hist[-1], "data_seconds": t_data,
                "train_seconds": t_train,
                "scf_convergence": "|dP|_F/2 <= 1e-5 and |dE| <= 1e-8 Eh, DIIS, 6-31G"})
    return res


def main(argv=None):
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="train and evaluate the learned SCF guess")
    ap.add_argument("--train", type=int, default=4096)
    ap.add_argument("--test", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--seed", type=int, default=2024)
    a = ap.parse_args(argv)

    def log(s):
        print(s, file=sys.stderr, flush=True)

    res = run(a.train, a.test, a.epochs, a.seed, log=log)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
