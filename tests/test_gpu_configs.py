"""The benchmark configurations beyond the headline, run as planned on the
GPU and checked against the oracle:

* batched: 384 float32 channels, lengths ~ U[20k, 200k] (seed 2026), fs =
  30 kHz, pink noise + biphasic spikes, Morlet sigma 6, 200 LOGSCALES
  10 Hz - 10 kHz, one Batch execute (about 67 GB of output in HBM);
* long signal: n = 2^22 AR(2) + 3-segment tone, sigma 4, 1024 LINFREQS over
  [0.001, 0.5] fs (34.4 GB of output);
* differentiation: all 64 signals of n = 16,384, 128 LOGSCALES over
  [0.01, 0.45] fs, loss sum w |W|^2, gradient through the autograd module.

Each is the reference's per-signal FCWT::cwt (src/fcwt/fcwt.cpp:447-529)
applied to every signal; the adjoint is its transpose.  Bar: rel-L2 <= 1e-5
per row (per gradient vector) against float64.
"""
import numpy as np
import pytest
import torch

import bench
import fcwt_amd
from oracle import oracle as O
from tests.gpu_util import run_device, tones_noise

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _band_class_rows(n, sc, fb):
    """One row per distinct padded band size M = 2^ceil(log2 E) of the plan,
    plus the first and last rows."""
    N = 1 << O.find2power(n)
    E, _, _ = O.rows_meta(N, sc, 1, True)
    first = {}
    for i, e in enumerate(E):
        first.setdefault(int(np.ceil(np.log2(max(int(e), 1)))), i)
    return sorted(set(first.values()) | {0, len(sc) - 1})


def test_batched_384_channels():
    fs, fn, fb = 30_000, 200, 6.0
    chans = bench.batched_channels(fs=fs)
    lengths = [c.size for c in chans]
    assert len(chans) == 384 and min(lengths) >= 20_000 and max(lengths) <= 200_000
    sc = fcwt_amd.Scales(fcwt_amd.Morlet(fb), fcwt_amd.FCWT_LOGSCALES, fs, 10.0, 10_000.0, fn).scales
    b = fcwt_amd.Batch(lengths, sc, fb)
    shp = b.shape()
    sizes = sorted({O.find2power(n) for n in lengths})
    assert shp["plans"] == len(sizes) == 4, (shp, sizes)
    total = sum(lengths)
    out = torch.empty(fn * total, dtype=torch.complex64, device="cuda")
    offs = np.concatenate([[0], np.cumsum([fn * n for n in lengths])[:-1]])
    xs = [torch.from_numpy(c).cuda() for c in chans]
    st = torch.cuda.current_stream().cuda_stream
    b.execute_device([x.data_ptr() for x in xs], [out.data_ptr() + 8 * int(o) for o in offs], st)
    torch.cuda.synchronize()
    b.check()
    flat = out.view(torch.float32)
    for k in range(0, flat.numel(), 1 << 28):
        assert torch.isfinite(flat[k: k + (1 << 28)]).all(), k
    # two channels of each padded size: the shortest and the longest
    pick = []
    for m in sizes:
        ids = [c for c, n in enumerate(lengths) if O.find2power(n) == m]
        pick += [min(ids, key=lambda c: lengths[c]), max(ids, key=lambda c: lengths[c])]
    for c in pick:
        n = lengths[c]
        got = out[int(offs[c]): int(offs[c]) + fn * n].view(fn, n).cpu().numpy()
        assert np.array_equal(got, run_device(chans[c], sc, fb)), c
        rows = _band_class_rows(n, sc, fb)
        ref = O.cwt_fp64(chans[c], sc, fb, rows=rows)
        err = O.rel_l2_rows(got[rows], ref)
        assert err.max() <= TOL, (c, n, rows[int(np.argmax(err))], err.max())


def _rows_rel_l2_device(a, b):
    """Per-row relative L2 of complex64 (fn, n) device tensors a vs b, in float64 sums."""
    out = []
    for r0 in range(0, a.shape[0], 64):
        d = (a[r0: r0 + 64] - b[r0: r0 + 64]).abs().double().pow(2).sum(dim=1)
        s = b[r0: r0 + 64].abs().double().pow(2).sum(dim=1)
        out.append((d / s).sqrt())
    return torch.cat(out).cpu().numpy()


def test_long_signal_2p22_x_1024():
    n, fs, fn, fb = 1 << 22, 1000, 1024, 4.0
    x1 = bench.long_signal(n, fs)
    x2 = tones_noise(n, fs=fs, fmax=450, seed=77)
    a, c = 0.5, -1.25
    x3 = (a * x1 + c * x2).astype(np.float32)
    sc = fcwt_amd.Scales(fcwt_amd.Morlet(fb), fcwt_amd.FCWT_LINFREQS, fs, 0.001 * fs, 0.5 * fs, fn).scales
    p = fcwt_amd.Plan(n, sc, fb)
    assert p.row_split()["two_pass"] > 0
    st = torch.cuda.current_stream().cuda_stream
    outs = []
    for x in (x1, x2, x3):
        dx = torch.from_numpy(x).cuda()
        o = torch.empty((fn, n), dtype=torch.complex64, device="cuda")
        p.execute_device(dx.data_ptr(), o.data_ptr(), st)
        outs.append(o)
    torch.cuda.synchronize()
    p.check()
    W1, W2, W3 = outs
    # sampled rows, one per band class, against the float64 oracle
    rows = _band_class_rows(n, sc, fb)
    got = W1[rows].cpu().numpy()
    ref = O.cwt_fp64(x1, sc, fb, rows=rows)
    err = O.rel_l2_rows(got, ref)
    assert err.max() <= TOL, (rows[int(np.argmax(err))], err.max())
    # every row: linearity W(a x1 + c x2) = a W(x1) + c W(x2) (x3 rounded to float32)
    lin = W1.mul_(a).add_(W2, alpha=c)
    del W2
    e = _rows_rel_l2_device(lin, W3)
    assert np.isfinite(e).all() and e.max() <= TOL, (int(np.argmax(e)), e.max())


def test_differentiation_all_64_signals():
    from fcwt_amd.autograd import CWT
    from tests.test_gpu_adjoint import _torch_reference_grad, rel

    B, n, fn, fs, fb = 64, 16_384, 128, 1000, 2.0
    x, w = bench.adjoint_signals(B, n, fn, fs)
    sc = O.scales(O.LOGSCALES, fs, 0.01 * fs, 0.45 * fs, fn)
    op = CWT(n, sc, fb)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    W = op(xt)
    loss = (torch.from_numpy(w).cuda() * W.abs() ** 2).sum()
    loss.backward()
    op.check()
    got = xt.grad.cpu().numpy()
    worst = 0.0
    for b in range(B):
        ref = _torch_reference_grad(x[b], w, sc, fb)
        worst = max(worst, rel(got[b], ref))
    assert worst <= TOL, worst
