"""GPU parity of the column path for wide bands (k_spec_transpose,
k_cols_step, k_cols_out, k_cols_out512): padded sizes N = 2^13 .. 2^22 with
bands up to N/2, against the float64 oracle, against the generic two-pass
kernels (FCWT_HIP_NO_COLS=1) on the same plan shape, and at N = 2^22 between
its two splits (512 x 2^13 and 256 x 2^14).

Bar: relative L2 <= 1e-5 per row (TOL).  At N = 2^22 the oracle checks a
sample of rows that spans every band size (a float64 4M-point transform per
row); the remaining rows are covered by the agreement with the two-pass
kernels and by linearity.
"""
import numpy as np
import pytest

import fcwt_amd
from oracle import oracle as O
from tests.gpu_util import run_device, tones_noise

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _rows_spanning_bands(n, sc, fb=2.0, k=10):
    """Row indices covering each distinct band size of the plan, at most k."""
    N = 1 << O.find2power(n)
    E, _, _ = O.rows_meta(N, sc, 1, True)
    sizes = {}
    for i, e in enumerate(E):
        sizes.setdefault(int(np.ceil(np.log2(max(int(e), 1)))), i)
    rows = sorted(set(sizes.values()) | {0, len(sc) - 1})
    if len(rows) > k:
        rows = sorted(set(rows[:: max(1, len(rows) // k)]) | {0, len(sc) - 1})
    return rows


@pytest.mark.parametrize("n,fs,fn", [(150_000, 1000, 40), (600_000, 1000, 32), (3_000_000, 1000, 24)])
def test_column_path_against_oracle(n, fs, fn, monkeypatch):
    """N = 2^18, 2^20, 2^22: LINFREQS 1..fs/2 with sigma 4 (the long-signal
    workload's scale family), so bands run from a few thousand bins to N/2
    (at N = 2^18 with the team path, which takes those rows by default,
    switched off)."""
    if n <= (1 << 18):
        monkeypatch.setenv("FCWT_HIP_NO_TEAMS", "1")
    x = tones_noise(n, fs=fs, fmax=0.45 * fs, seed=n)
    sc = O.scales(O.LINFREQS, fs, 1, fs / 2, fn)
    p = fcwt_amd.Plan(n, sc, 4.0)
    split = p.row_split()
    assert split["two_pass"] > 0 and split["team"] == 0, split
    p.profile(True)
    got = run_device(x, sc, 4.0, plan=p)
    kinds = p.profile_read()
    p.profile(False)
    assert {"row_columns", "column_transpose"} <= set(kinds) and "row_spreader" not in kinds, kinds
    assert np.all(np.isfinite(got))
    rows = _rows_spanning_bands(n, sc, 4.0)
    ref = O.cwt_fp64(x, sc, 4.0, rows=rows)
    err = O.rel_l2_rows(got[rows], ref)
    assert err.max() <= TOL, f"row {rows[int(np.argmax(err))]}: rel-L2 {err.max():.3e}"


@pytest.mark.parametrize("n", [5000, 8192, 10_000, 16_384, 20_000, 50_000, 100_000, 200_000])
def test_column_path_every_row(n, monkeypatch):
    """N = 2^13 .. 2^18 (the team path off where it would apply): every row
    against the oracle."""
    monkeypatch.setenv("FCWT_HIP_NO_TEAMS", "1")
    x = tones_noise(n, seed=n + 1)
    sc = O.scales(O.LOGSCALES, 1000, 0.5, 499, 48)
    p = fcwt_amd.Plan(n, sc, 2.0)
    assert p.row_split()["two_pass"] > 0
    got = run_device(x, sc, 2.0, plan=p)
    ref = O.cwt_fp64(x, sc, 2.0, 1, True, True)
    err = O.rel_l2_rows(got, ref)
    assert err.max() <= TOL, f"row {int(np.argmax(err))}: rel-L2 {err.max():.3e}"


@pytest.mark.parametrize("threads,flags", [(6, 0), (96, 0), (6, 1)])
def test_column_path_stale_segments(threads, flags, monkeypatch):
    """Thread counts that skip bins near Nyquist: the wide rows carry earlier
    scales' products as extra band segments (window evaluated per segment)."""
    monkeypatch.setenv("FCWT_HIP_NO_TEAMS", "1")
    x = tones_noise(40_000, seed=4)
    sc = O.scales(O.LOGSCALES, 1000, 1, 480, 64)
    got = run_device(x, sc, 2.0, threads=threads, flags=flags)
    ref = O.cwt_fp64(x, sc, 2.0, threads, not (flags & 1), True)
    err = O.rel_l2_rows(got, ref)
    nz = np.linalg.norm(ref, axis=1) > 0
    assert np.all(got[~nz] == 0)
    assert err[nz].max() <= TOL, err.max()


def test_column_path_low_frequency_rows(monkeypatch):
    """fs = 100, LOGSCALES 0.01..10 Hz, 8 threads, n = 30,000: rows with no
    coverage of their own repeat earlier products."""
    monkeypatch.setenv("FCWT_HIP_NO_TEAMS", "1")
    x = tones_noise(30_000, fs=100, fmax=31, seed=5)
    sc = O.scales(O.LOGSCALES, 100, 0.01, 10, 200)
    got = run_device(x, sc, 2.0, threads=8)
    ref = O.cwt_fp64(x, sc, 2.0, 8, True, True)
    err = O.rel_l2_rows(got, ref)
    nz = np.linalg.norm(ref, axis=1) > 0
    assert err[nz].max() <= TOL, err.max()


@pytest.mark.parametrize("n", [10_000, 150_000, 1_500_000])
def test_column_path_matches_two_pass_kernels(n, monkeypatch):
    """The same plan through the generic two-pass kernels: same rows, the
    float32 results agree to float32 rounding (the team path, which takes
    the wide rows at N = 2^18 by default, switched off)."""
    monkeypatch.setenv("FCWT_HIP_NO_TEAMS", "1")
    x = tones_noise(n, seed=7)
    sc = O.scales(O.LINFREQS, 1000, 1, 500, 16)
    cols = run_device(x, sc, 4.0)
    monkeypatch.setenv("FCWT_HIP_NO_COLS", "1")
    p = fcwt_amd.Plan(n, sc, 4.0)
    assert p.row_split()["two_pass"] > 0
    two = run_device(x, sc, 4.0, plan=p)
    assert O.rel_l2_rows(cols, two).max() <= 2e-6


def test_column_path_linearity_full_rows():
    """n = 2^22, 1024 LINFREQS rows would not fit a test; 64 rows over the
    same range: cwt(a + b) = cwt(a) + cwt(b) on every row."""
    n = 1 << 22
    a = tones_noise(n, seed=11)
    b = tones_noise(n, seed=12, fmax=50)
    sc = O.scales(O.LINFREQS, 1000, 1, 500, 64)
    p = fcwt_amd.Plan(n, sc, 4.0)
    ya = run_device(a, sc, 4.0, plan=p)
    yb = run_device(b, sc, 4.0, plan=p)
    yab = run_device(a + b, sc, 4.0, plan=p)
    assert O.rel_l2_rows(yab, ya + yb).max() < 3e-5


@pytest.mark.parametrize("n,threads", [((1 << 21) + 12_345, 1), (1 << 22, 6)])
def test_column_path_splits_at_2_22(n, threads, monkeypatch):
    """N = 2^22 splits as 512 x 2^13 (the default) or, with
    FCWT_HIP_COLS_B256=1, as 256 x 2^14: sampled rows of both against the
    oracle, every row of the two against each other.  threads = 6 skips bins
    near Nyquist, so the wide rows carry several band segments."""
    x = tones_noise(n, seed=n + threads)
    sc = O.scales(O.LINFREQS, 1000, 1, 500, 20)
    got = {}
    for split in ("512", "256"):
        if split == "256":
            monkeypatch.setenv("FCWT_HIP_COLS_B256", "1")
        p = fcwt_amd.Plan(n, sc, 4.0, threads)
        assert p.row_split()["two_pass"] > 0
        got[split] = run_device(x, sc, 4.0, threads=threads, plan=p)
        p.close()
    assert np.all(np.isfinite(got["512"]))
    rows = [0, 5, 12, 19]
    ref = O.cwt_fp64(x, sc, 4.0, threads, True, True, rows=rows)
    for split, y in got.items():
        err = O.rel_l2_rows(y[rows], ref)
        assert err.max() <= TOL, f"{split}: row {rows[int(np.argmax(err))]}: rel-L2 {err.max():.3e}"
    nz = np.linalg.norm(got["256"], axis=1) > 0
    assert np.all(got["512"][~nz] == 0)
    assert O.rel_l2_rows(got["512"][nz], got["256"][nz]).max() <= 2e-6


def test_column_path_two_signals_at_2_22():
    """Two signals of N = 2^22 in one batch plan (the transposed spectra and
    scratch slots of both, pass 2 writing each signal's rows): each signal's
    output equals its one-signal plan's bit for bit."""
    import torch

    lengths = [(1 << 21) + 1, (1 << 22) - 3]
    xs = [tones_noise(n, seed=k + 30) for k, n in enumerate(lengths)]
    scs = [O.scales(O.LINFREQS, 1000, 1, 500, 6), O.scales(O.LINFREQS, 1000, 20, 490, 5)]
    b = fcwt_amd.Batch(lengths, scs, 4.0)
    assert b.shape()["plans"] == 1
    dx = [torch.from_numpy(x).cuda() for x in xs]
    outs = [torch.empty((s.size, x.size), dtype=torch.complex64, device="cuda") for s, x in zip(scs, xs)]
    b.execute_device([d.data_ptr() for d in dx], [o.data_ptr() for o in outs],
                     torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    b.check()
    for x, sc, o in zip(xs, scs, outs):
        assert np.array_equal(o.cpu().numpy(), run_device(x, sc, 4.0)), x.size


def test_tuning_file_routes_wide_rows(tmp_path):
    """A tuning file that names the column path for N = 2^17 (what
    fcwt_hip_tune writes when that path times faster there) moves the wide
    rows off the XCD-team kernel; the result stays within the bar.  The
    built-in routing is restored afterwards (the tuned table is process-wide)."""
    from fcwt_amd import _native as nat

    def load(path_choice):
        f = tmp_path / f"n131072_t1_{path_choice}.fcwt_hip"
        f.write_text("fcwt_hip tuning 1\narch gfx950\nN 131072\nthreads 1\neffort 1\nteam_buffers 3\n"
                     f"cols_mb 128\npath {path_choice}\ntime_ms 1\n")
        nat.check(nat.lib().fcwt_hip_tune_load(str(f).encode()), "fcwt_hip_tune_load")

    x = tones_noise(100_000, seed=21)
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 64)
    load("cols")
    try:
        p = fcwt_amd.Plan(x.size, sc, 2.0)
        split = p.row_split()
        assert split["team"] == 0 and split["two_pass"] > 0, split
        got = run_device(x, sc, 2.0, plan=p)
    finally:
        load("default")
    assert fcwt_amd.Plan(x.size, sc, 2.0).row_split()["team"] > 0
    ref = O.cwt_fp64(x, sc, 2.0)
    assert O.rel_l2_rows(got, ref).max() <= TOL
