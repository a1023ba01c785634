"""MLP kernels of the learned SCF guess (run on an MI355X) against a torch
FP32 CPU reference that rounds to BF16 at the same points as the HIP path,
and the trained guess against the core-Hamiltonian guess."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _lib():
    from librint_amd._bindings import library
    return library


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _bf(x):
    """Round an FP32 CPU tensor to BF16 values (kept as FP32)."""
    import torch
    return x.to(torch.bfloat16).to(torch.float32)


@pytest.mark.parametrize("M,N,K,ta,tb", [(16, 16, 16, 0, 0), (37, 53, 29, 0, 0), (128, 256, 48, 0, 0),
                                         (1000, 16, 256, 0, 0), (256, 48, 1000, 1, 0), (300, 256, 16, 0, 1),
                                         (65, 70, 33, 1, 1)])
def test_gemm_bf16_vs_fp32_reference(gpu, M, N, K, ta, tb):
    import torch
    _say("gemm", M, N, K, ta, tb)
    g = torch.Generator().manual_seed(M * 7 + N)
    A = torch.randn(K, M, generator=g) if ta else torch.randn(M, K, generator=g)
    B = torch.randn(N, K, generator=g) if tb else torch.randn(K, N, generator=g)
    bias = torch.randn(N, generator=g)
    Ab, Bb = A.to(torch.bfloat16), B.to(torch.bfloat16)
    opA = Ab.float().T if ta else Ab.float()
    opB = Bb.float().T if tb else Bb.float()
    for act in (0, 1):
        ref = opA @ opB + bias
        if act:
            ref = torch.relu(ref)
        dA, dB, db = Ab.to(gpu), Bb.to(gpu), bias.to(gpu)
        C = torch.full((M, N), 7.0, device=gpu)
        Cb = torch.empty((M, N), dtype=torch.bfloat16, device=gpu)
        lda = M if ta else K
        ldb = K if tb else N
        rc = _lib().librint_mlp_gemm(M, N, K, _p(dA), lda, ta, _p(dB), ldb, tb, _p(C), N, _p(db), act, 0, _p(Cb), N,
                                     None)
        assert rc == 0
        torch.cuda.synchronize()
        got = C.cpu()
        scale = (opA.abs() @ opB.abs()).max().item() + 1.0
        assert (got - ref).abs().max().item() < 1e-5 * scale
        np.testing.assert_array_equal(Cb.cpu().float().numpy(), got.to(torch.bfloat16).float().numpy())
    # accumulate into C
    C0 = torch.randn(M, N, generator=g)
    C = C0.to(gpu)
    assert _lib().librint_mlp_gemm(M, N, K, _p(dA), lda, ta, _p(dB), ldb, tb, _p(C), N, None, 0, 1, None, 0, None) == 0
    torch.cuda.synchronize()
    ref = C0 + opA @ opB
    assert (C.cpu() - ref).abs().max().item() < 1e-5 * scale


def test_train_step_gradients_vs_torch(gpu):
    """Forward, loss, backward of the HIP MLP against torch autograd on the
    CPU in FP32 with the same BF16 rounding of weights, activations and
    output gradients."""
    import torch
    from librint_amd.guess import MLP
    net = MLP(seed=3, dims=(48, 64, 64, 64, 16))
    rows = 300
    g = torch.Generator().manual_seed(5)
    X = torch.randn(rows, 48, generator=g)
    T = torch.randn(rows, 2, generator=g)
    W = (torch.rand(rows, 2, generator=g) > 0.3).float()
    xb = net.to_bf16(X.to(gpu))
    y = net.forward(xb, keep=True)
    scale = float(W.sum())
    loss, dy_f, dy_b = net.loss_grad(y, T.to(gpu), W.to(gpu), scale)
    dW, db = net.backward(dy_f, dy_b)
    torch.cuda.synchronize()
    # reference: BF16 weights, BF16 layer inputs, BF16 pre-activation
    # gradients into the GEMMs (as the HIP path), FP32 everywhere else
    Ws = [_bf(w.cpu()) for w in net.W]
    bs = [b.cpu() for b in net.b]
    h = _bf(X)
    hs, zs = [h], []
    for i in range(4):
        z = h @ Ws[i] + bs[i]
        zs.append(z)
        if i < 3:
            h = _bf(torch.relu(z))
            hs.append(h)
    yref = zs[-1]
    assert (y.cpu() - yref).abs().max().item() < 1e-4 * (yref.abs().max().item() + 1)
    d = yref[:, :2] - T
    lref = (W * d * d).sum() / scale
    assert abs(loss.item() - lref.item()) < 1e-4 * abs(lref.item())
    dy = torch.zeros_like(yref)
    dy[:, :2] = 2 * W * d / scale
    dpre = dy
    for i in reversed(range(4)):
        dWref = hs[i].T @ _bf(dpre)
        dbref = dpre.sum(0)
        s = dWref.abs().max().item() + 1e-12
        assert (dW[i].cpu() - dWref).abs().max().item() < 2e-3 * s, i
        assert (db[i].cpu() - dbref).abs().max().item() < 2e-3 * (dbref.abs().max().item() + 1e-12), i
        if i:
            dz = _bf(dpre) @ Ws[i].T
            dpre = dz * (zs[i - 1] > 0).float()


def test_adam_and_conversion(gpu):
    import torch
    n = 1000
    g = torch.Generator().manual_seed(9)
    w0, gr = torch.randn(n, generator=g), torch.randn(n, generator=g)
    m0, v0 = torch.randn(n, generator=g) * 0.1, torch.rand(n, generator=g) * 0.1
    w, m, v = w0.to(gpu), m0.to(gpu), v0.to(gpu)
    wb = torch.empty(n, dtype=torch.bfloat16, device=gpu)
    lr, b1, b2, eps, step = 1e-2, 0.9, 0.999, 1e-8, 3
    assert _lib().librint_mlp_adam(n, _p(w), _p(gr.to(gpu)), _p(m), _p(v), lr, b1, b2, eps, step, _p(wb), None) == 0
    torch.cuda.synchronize()
    mr = b1 * m0 + (1 - b1) * gr
    vr = b2 * v0 + (1 - b2) * gr * gr
    wr = w0 - lr * (mr / (1 - b1 ** step)) / (torch.sqrt(vr / (1 - b2 ** step)) + eps)
    np.testing.assert_allclose(w.cpu().numpy(), wr.numpy(), rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(wb.cpu().float().numpy(), w.cpu().to(torch.bfloat16).float().numpy())
    x = torch.randn(777, generator=g).to(gpu)
    y = torch.empty(777, dtype=torch.bfloat16, device=gpu)
    assert _lib().librint_mlp_to_bf16(777, _p(x), _p(y), None) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().float().numpy(), x.cpu().to(torch.bfloat16).float().numpy())
    assert _lib().librint_mlp_adam(n, _p(w), _p(gr.to(gpu)), _p(m), _p(v), lr, b1, b2, eps, 0, None, None) == -1


def test_trained_guess_cuts_iterations_same_energy(gpu):
    """The learned-guess bar at reduced size: a network trained (on the GPU,
    by the library's own BF16 MFMA kernels) on 512 seeded perturbed C2-C6
    alkanes, whose converged densities come from the GPU SCF, against the
    core-Hamiltonian guess on 64 held-out C2-C6 alkanes.  It must cut the SCF
    iterations by at least 30 % and leave every converged energy within
    1e-10 Eh of the core-Hamiltonian-guess one."""
    from librint_amd import guess as G
    res = G.run(512, 64, epochs=20, seed=2024, log=_say)
    _say(res)
    assert res["final_train_mse"] < 0.2 * res["first_train_mse"]
    assert res["converged_all"]
    assert res["molecules"] == 64
    assert res["max_energy_difference_Eh"] <= 1e-10
    assert res["reduction"] >= 0.30
