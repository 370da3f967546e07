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


def _sizes_batch(seed=9):
    """Channels of every padded size the XCD-team kernel takes (2^15 .. 2^18),
    sigma 6 LOGSCALES 10 Hz .. 10 kHz at fs 30 kHz: wide rows on the team
    kernel and direct rows of every band from 2^1 up."""
    lengths = [17_000, 30_011, 33_000, 60_000, 65_536, 90_001, 131_072, 140_000, 200_000]
    xs = [tones_noise(n, fs=30_000, seed=n, fmax=12_000) for n in lengths]
    sc = O.scales(O.LOGSCALES, 30_000, 10, 10_000, 40)
    return lengths, xs, sc


def test_one_team_launch_for_every_padded_size(monkeypatch):
    """A batch whose signals pad to 2^15, 2^16, 2^17 and 2^18 makes one
    row-kernel launch per execute (k_rows_team_multi: the team rows of the four
    plans, then their direct rows); each signal's output equals its
    one-signal plan's and the per-size team launches'
    (FCWT_HIP_BATCH_TEAM_LAUNCHES=separate) bit for bit; sampled rows of every
    size against the oracle."""
    lengths, xs, sc = _sizes_batch()
    b = fcwt_amd.Batch(lengths, sc, 6.0)
    shp = b.shape()
    assert shp["plans"] == 4 and shp["team_rows"] > 0 and shp["row_launches"] == 1, shp
    got = _run_batch(b, xs, [sc.size] * len(xs))
    again = _run_batch(b, xs, [sc.size] * len(xs))  # counters cleared before every launch
    monkeypatch.setenv("FCWT_HIP_BATCH_TEAM_LAUNCHES", "separate")
    sep = fcwt_amd.Batch(lengths, sc, 6.0)
    assert sep.shape()["row_launches"] > 1
    ref_sep = _run_batch(sep, xs, [sc.size] * len(xs))
    for k, x in enumerate(xs):
        assert np.array_equal(got[k], again[k]), x.size
        assert np.array_equal(got[k], ref_sep[k]), x.size
        assert np.array_equal(got[k], run_device(x, sc, 6.0)), x.size
    N = [1 << O.find2power(n) for n in lengths]
    rows = [0, 5, 12, 20, 28, 39]
    for k in (0, 3, 5, 8):
        ref = O.cwt_fp64(xs[k], sc, 6.0, rows=rows)
        err = O.rel_l2_rows(got[k][rows], ref)
        assert err.max() <= TOL, (lengths[k], N[k], rows[int(np.argmax(err))], err.max())


def test_one_team_launch_failure_is_reported_by_every_plan(monkeypatch):
    """A spin limit of zero polls fails the launch's first team wait: every
    plan of the batch reports the failure, every team row of the launch is
    NaN, and the next execute (normal spin) is correct."""
    from fcwt_amd import _native as nat

    lengths, xs, sc = _sizes_batch(seed=10)
    monkeypatch.setenv("FCWT_HIP_TEAM_SPIN", "0")
    b = fcwt_amd.Batch(lengths, sc, 6.0)
    assert b.shape()["row_launches"] == 1
    st = torch.cuda.current_stream()
    dx = [torch.from_numpy(x).cuda() for x in xs]
    outs = [torch.zeros((sc.size, x.size), dtype=torch.complex64, device="cuda") for x in xs]
    b.execute_device([d.data_ptr() for d in dx], [o.data_ptr() for o in outs], st.cuda_stream)
    torch.cuda.synchronize()
    with pytest.raises(nat.FcwtError):
        b.check()
    b.check()  # reported once
    assert all(torch.isnan(o).any().item() for o in outs)  # every size's team rows
    monkeypatch.delenv("FCWT_HIP_TEAM_SPIN")
    got = _run_batch(b, xs, [sc.size] * len(xs))
    for k in (0, 4, 8):
        assert np.array_equal(got[k], run_device(xs[k], sc, 6.0)), lengths[k]


def test_team_launch_beside_column_path_and_graph_replays():
    """Signals padded to 2^16, 2^18 and 2^20: the first two share one team
    launch, the 2^20 signal's wide rows take the column path in launches of
    their own; executes with the same buffers replay the captured graph.
    Every signal equals its one-signal plan bit for bit, on every replay."""
    lengths = [40_000, 150_000, 600_000]
    xs = [tones_noise(n, fs=30_000, seed=n + 5, fmax=12_000) for n in lengths]
    sc = O.scales(O.LOGSCALES, 30_000, 10, 10_000, 24)
    b = fcwt_amd.Batch(lengths, sc, 6.0)
    shp = b.shape()
    assert shp["plans"] == 3 and shp["row_launches"] > 1, shp
    refs = [run_device(x, sc, 6.0) for x in xs]
    st = torch.cuda.current_stream()
    dx = [torch.from_numpy(x).cuda() for x in xs]
    outs = [torch.empty((sc.size, x.size), dtype=torch.complex64, device="cuda") for x in xs]
    for k in range(4):  # eager, capture, two replays
        for o in outs:
            o.zero_()
        b.execute_device([d.data_ptr() for d in dx], [o.data_ptr() for o in outs], st.cuda_stream)
        torch.cuda.synchronize()
        b.check()
        for o, r, n in zip(outs, refs, lengths):
            assert np.array_equal(o.cpu().numpy(), r), (k, n)
    assert b.graph_launches() == 3
