"""GPU adjoint (fcwt_hip_adjoint_device) and the torch.autograd.Function on
top of it, against the float64 oracle adjoint (itself pinned in
test_oracle.py by the explicit transpose and finite differences).

Bar: relative L2 <= 1e-5 of the gradient vector (TOL).  The differentiation
workload: 64 signals of n = 16,384 (N(0,1) plus random chirps), Morlet
sigma = 2, 128 LOGSCALES over [0.01, 0.45] x fs, loss sum w |W|^2 with
w ~ U[0,1] (fn x n, seeded); its gradient is checked against float64
central differences (h = 1e-3) at 32 random samples and against a
torch.fft-composed float64 reference gradient.
"""
import numpy as np
import pytest
import torch

import fcwt_amd
from fcwt_amd.autograd import CWT
from oracle import oracle as O
from tests.gpu_util import chirp_noise, tones_noise

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _adjoint_device(plan, G, complex_input=False):
    g = torch.from_numpy(np.ascontiguousarray(G, np.complex64)).cuda()
    xbar = torch.empty(plan.n, dtype=torch.complex64 if complex_input else torch.float32, device="cuda")
    plan.adjoint_device(g.data_ptr(), xbar.data_ptr(), torch.cuda.current_stream().cuda_stream, complex_input)
    torch.cuda.synchronize()
    return xbar.cpu().numpy()


@pytest.mark.parametrize("n,f0,f1,fn", [(1000, 1, 450, 40), (10_000, 1, 101, 300), (16_384, 10, 450, 128),
                                         (50_000, 0.5, 499, 48), (100_000, 1, 450, 200)])
def test_adjoint_matches_oracle(n, f0, f1, fn):
    """Direct-row, two-pass and four-step forward-FFT sizes."""
    rng = np.random.default_rng(n)
    sc = O.scales(O.LOGSCALES, 1000, f0, f1, fn)
    G = (rng.standard_normal((fn, n)) + 1j * rng.standard_normal((fn, n))).astype(np.complex64)
    p = fcwt_amd.Plan(n, sc, 2.0)
    got = _adjoint_device(p, G)
    ref = O.adjoint_fp64(G, sc, 2.0)
    assert rel(got, ref) <= TOL, rel(got, ref)


def test_adjoint_complex_input_and_host_api():
    rng = np.random.default_rng(3)
    n, fn = 7000, 60
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, fn)
    G = (rng.standard_normal((fn, n)) + 1j * rng.standard_normal((fn, n))).astype(np.complex64)
    p = fcwt_amd.Plan(n, sc, 2.0, normalize=False)
    got = _adjoint_device(p, G, complex_input=True)
    ref = O.adjoint_fp64(G, sc, 2.0, normalize=False, complex_input=True)
    assert rel(got, ref) <= TOL
    host = p.adjoint_host(G, complex_input=True)
    assert np.array_equal(host, got)
    assert np.array_equal(p.adjoint_host(G), _adjoint_device(p, G))


def test_dot_product_identity_on_gpu():
    """Re <W(x), G> = <x, xbar> with both sides from the GPU."""
    rng = np.random.default_rng(11)
    n, fn = 100_000, 3000
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, fn)
    x = tones_noise(n, seed=5)
    p = fcwt_amd.Plan(n, sc, 2.0)
    from tests.gpu_util import run_device

    W = run_device(x, sc, 2.0, plan=p)
    G = (rng.standard_normal((fn, n)) + 1j * rng.standard_normal((fn, n))).astype(np.complex64)
    xbar = _adjoint_device(p, G)
    lhs = np.real(np.vdot(W.astype(np.complex128).ravel(), G.astype(np.complex128).ravel()))
    rhs = float(np.dot(x.astype(np.float64), xbar.astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs), (lhs, rhs)


def _c4_batch(B, n=16_384, fn=128, fs=1000, seed=2024):
    rng = np.random.default_rng(seed)
    t = np.arange(n) / fs
    xs = []
    for b in range(B):
        f0, f1 = rng.uniform(5, 450, 2)
        chirp = np.sin(2 * np.pi * (f0 + (f1 - f0) * t / (2 * t[-1])) * t)
        xs.append(rng.standard_normal(n) + chirp)
    x = np.stack(xs).astype(np.float32)
    w = rng.uniform(0, 1, (fn, n)).astype(np.float32)
    sc = O.scales(O.LOGSCALES, fs, 0.01 * fs, 0.45 * fs, fn)
    return x, w, sc


def _torch_reference_grad(x, w, sc, fb=2.0):
    """float64 gradient through torch.fft of the same linear map (effective
    windows from the oracle), CPU."""
    n = x.shape[-1]
    N = 1 << O.find2power(n)
    wins = torch.from_numpy(np.stack([w_.copy() for _, w_ in O.effective_windows(sc, fb, N)]))
    xt = torch.from_numpy(x.astype(np.float64)).requires_grad_(True)
    X = torch.fft.rfft(torch.nn.functional.pad(xt, (0, N - n)))[..., : N // 2]
    Y = X[..., None, :] * wins
    full = torch.nn.functional.pad(Y, (0, N - N // 2))
    W = torch.fft.ifft(full)[..., :n]  # = (1/N) * unnormalised backward DFT
    loss = (torch.from_numpy(w.astype(np.float64)) * W.abs() ** 2).sum()
    loss.backward()
    return xt.grad.numpy()


def test_autograd_c4_against_torch_fft_reference():
    x, w, sc = _c4_batch(B=4)
    op = CWT(x.shape[-1], sc, 2.0)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    W = op(xt)
    assert W.shape == (4, 128, x.shape[-1])
    loss = (torch.from_numpy(w).cuda() * W.abs() ** 2).sum()
    loss.backward()
    got = xt.grad.cpu().numpy()
    ref = _torch_reference_grad(x, w, sc)
    assert rel(got, ref) <= TOL, rel(got, ref)


def test_autograd_finite_differences():
    """float64 central differences (h = 1e-3) at 32 random samples of one C4 signal."""
    x, w, sc = _c4_batch(B=1, seed=7)
    x = x[0]
    n = x.size
    op = CWT(n, sc, 2.0)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    loss = (torch.from_numpy(w).cuda() * op(xt).abs() ** 2).sum()
    loss.backward()
    got = xt.grad.cpu().numpy().astype(np.float64)
    N = 1 << O.find2power(n)
    wins = np.stack([w_.copy() for _, w_ in O.effective_windows(sc, 2.0, N)])

    def L(v):
        X = np.fft.rfft(np.pad(v, (0, N - n)))[: N // 2]
        full = np.zeros((len(sc), N), np.complex128)
        full[:, : N // 2] = X * wins
        W = np.fft.ifft(full, axis=1)[:, :n]
        return float((w * np.abs(W) ** 2).sum())

    rng = np.random.default_rng(1)
    h = 1e-3
    x64 = x.astype(np.float64)
    scale = np.abs(got).max()
    for t in rng.choice(n, 32, replace=False):
        e = np.zeros(n)
        e[t] = h
        fd = (L(x64 + e) - L(x64 - e)) / (2 * h)
        assert abs(fd - got[t]) <= 1e-4 * scale, (t, fd, got[t])


def test_autograd_complex_input():
    rng = np.random.default_rng(4)
    n, fn = 3000, 20
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, fn)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    op = CWT(n, sc, 2.0)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    W = op(xt)
    G = torch.randn(W.shape, dtype=torch.complex64, device="cuda")
    (W * G.conj()).real.sum().backward()  # dL/dW (PyTorch convention) = G
    ref = O.adjoint_fp64(G.cpu().numpy(), sc, 2.0, complex_input=True)
    assert rel(xt.grad.cpu().numpy(), ref) <= TOL


def test_batched_adjoint_and_forward_match_per_signal():
    """fcwt_hip_adjoint_batch_device over 5 signals equals 5 one-signal
    adjoints, and the module's batched forward equals per-signal forwards."""
    rng = np.random.default_rng(8)
    B, n, fn = 5, 20_000, 50
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, fn)
    op = CWT(n, sc, 2.0)
    x = torch.from_numpy(rng.standard_normal((B, n)).astype(np.float32)).cuda()
    W = op(x)
    from tests.gpu_util import run_device

    for b in (0, 4):
        assert np.array_equal(W[b].cpu().numpy(), run_device(x[b].cpu().numpy(), sc, 2.0, plan=op.plan))
    G = torch.randn((B, fn, n), dtype=torch.complex64, device="cuda")
    xbar = torch.empty((B, n), dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    op.plan.adjoint_batch_device(G.data_ptr(), xbar.data_ptr(), B, st)
    torch.cuda.synchronize()
    for b in range(B):
        one = _adjoint_device(op.plan, G[b].cpu().numpy())
        assert np.array_equal(xbar[b].cpu().numpy(), one), b
