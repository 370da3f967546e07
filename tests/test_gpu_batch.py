"""GPU batches (fcwt_hip_batch_*): many signals of uneven length, each with
its own scales, through merged launches.  Each signal's output must equal,
bit for bit, what a one-signal plan gives for it (same kernels, same
arithmetic), and sampled rows must meet the oracle bar (rel-L2 <= 1e-5).
"""
import numpy as np
import pytest
import torch

import fcwt_amd
from oracle import oracle as O
from tests.gpu_util import run_device, tones_noise

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _run_batch(batch, xs, fns, stream=None):
    st = stream or torch.cuda.current_stream()
    dx = [torch.from_numpy(x).cuda() for x in xs]
    outs = [torch.empty((f, x.size), dtype=torch.complex64, device="cuda") for f, x in zip(fns, xs)]
    batch.execute_device([d.data_ptr() for d in dx], [o.data_ptr() for o in outs], st.cuda_stream)
    torch.cuda.synchronize()
    batch.check()
    return [o.cpu().numpy() for o in outs]


def test_uneven_lengths_and_scales_match_single_plans():
    rng = np.random.default_rng(42)
    lengths = [1, 7, 1000, 4097, 20_000, 30_011, 60_000, 65_536, 100_000, 131_072, 150_001, 200_000]
    xs = [tones_noise(n, seed=n) for n in lengths]
    scs = []
    for n in lengths:
        fn = int(rng.integers(8, 40))
        f0 = float(rng.uniform(0.5, 5))
        scs.append(O.scales(O.LOGSCALES, 1000, f0, 480, fn))
    b = fcwt_amd.Batch(lengths, scs, 2.0)
    assert b.shape()["plans"] == len({O.find2power(n) for n in lengths})
    got = _run_batch(b, xs, [s.size for s in scs])
    for x, sc, g in zip(xs, scs, got):
        ref = run_device(x, sc, 2.0)
        assert np.array_equal(g, ref), x.size
    for i in (2, 5, 8, 11):
        rows = [0, scs[i].size // 2, scs[i].size - 1]
        ref = O.cwt_fp64(xs[i], scs[i], 2.0, rows=rows)
        assert O.rel_l2_rows(got[i][rows], ref).max() <= TOL


def test_many_channels_one_size_merge_team_rows():
    """48 channels of N = 2^17 in one plan: their wide rows share one team launch."""
    lengths = [100_000 + 600 * k for k in range(48)]
    sc = O.scales(O.LOGSCALES, 30_000, 10, 10_000, 60)
    xs = [tones_noise(n, fs=30_000, seed=k, fmax=12_000) for k, n in enumerate(lengths)]
    b = fcwt_amd.Batch(lengths, sc, 6.0)
    shp = b.shape()
    assert shp["plans"] == 1 and shp["team_rows"] > 0
    got = _run_batch(b, xs, [60] * 48)
    for k in (0, 17, 47):
        assert np.array_equal(got[k], run_device(xs[k], sc, 6.0)), k
    ref = O.cwt_fp64(xs[47], sc, 6.0, rows=[0, 1, 30, 59])
    assert O.rel_l2_rows(got[47][[0, 1, 30, 59]], ref).max() <= TOL


def test_new_output_buffers_between_calls():
    lengths = [5000, 5003, 9000]
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 20)
    xs = [tones_noise(n, seed=3 + n) for n in lengths]
    b = fcwt_amd.Batch(lengths, sc, 2.0)
    first = _run_batch(b, xs, [20] * 3)
    second = _run_batch(b, [x[::-1].copy() for x in xs], [20] * 3)
    for x, f, s in zip(xs, first, second):
        assert np.array_equal(f, run_device(x, sc, 2.0))
        assert np.array_equal(s, run_device(x[::-1].copy(), sc, 2.0))


def test_graph_replays_match_eager_execution(monkeypatch):
    """Executes with the same buffers replay a captured HIP graph.  Every
    replay, with eager work and synchronisations between them, equals the
    eager result bit for bit; other output buffers drop the graph and stay
    correct, and coming back captures again."""
    lengths = [30_000, 70_000, 150_000, 160_000]
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 48)
    xs = [tones_noise(n, seed=n) for n in lengths]
    monkeypatch.setenv("FCWT_HIP_NO_GRAPHS", "1")
    eager = fcwt_amd.Batch(lengths, sc, 2.0)
    ref = _run_batch(eager, xs, [48] * 4)
    assert eager.graph_launches() == 0
    monkeypatch.delenv("FCWT_HIP_NO_GRAPHS")
    b = fcwt_amd.Batch(lengths, sc, 2.0)
    dx = [torch.from_numpy(x).cuda() for x in xs]
    out_a = [torch.empty((48, n), dtype=torch.complex64, device="cuda") for n in lengths]
    out_b = [torch.empty((48, n), dtype=torch.complex64, device="cuda") for n in lengths]
    st = torch.cuda.current_stream()

    def run(outs):
        for o in outs:
            o.zero_()
        b.execute_device([d.data_ptr() for d in dx], [o.data_ptr() for o in outs], st.cuda_stream)
        torch.cuda.synchronize()
        b.check()
        return [o.cpu().numpy() for o in outs]

    for k in range(5):
        for g, r in zip(run(out_a), ref):
            assert np.array_equal(g, r), k
        # the program's own eager work between replays
        assert torch.isfinite(torch.randn(1 << 20, device="cuda").cumsum(0)).all()
    assert b.graph_launches() == 4  # the first execute is eager, the second captures
    for g, r in zip(run(out_b), ref):
        assert np.array_equal(g, r)
    for k in range(3):
        for g, r in zip(run(out_a), ref):
            assert np.array_equal(g, r), k
    assert b.graph_launches() == 6
