"""GPU parity: the HIP path against the oracle (float64 DFT of the same padded
problem, with the reference's float32 window values and O1 reuse).

Bar: relative L2 error <= 1e-5 per output row (TOL below), float32 arithmetic
on the GPU.  Each case also checks that the GPU's error is of the same order
as a float32 CPU evaluation of the same path (the oracle's float32 port).
"""
import numpy as np
import pytest

import fcwt_amd
from oracle import oracle as O
from tests.gpu_util import chirp_noise, run_device, tones_noise

pytestmark = pytest.mark.gpu

TOL = 1e-5  # per-row relative L2 vs the float64 oracle


def _check(x, sc, fb, threads=1, normalize=True, flags=0, rows=None, tol=TOL):
    got = run_device(x, sc, fb, threads, normalize, flags)
    avx = not (flags & 1)
    sel = list(range(len(sc))) if rows is None else list(rows)
    ref = O.cwt_fp64(x, sc, fb, threads, avx, normalize, rows=sel)
    err = O.rel_l2_rows(got[sel], ref)
    nz = np.linalg.norm(ref, axis=1) > 0
    assert np.all(np.isfinite(got))
    assert np.all(got[sel][~nz] == 0), "rows the reference leaves all-zero must be zero"
    worst = int(np.argmax(np.where(nz, err, 0))) if nz.any() else 0
    assert err[nz].max(initial=0) <= tol, f"row {sel[worst]}: rel-L2 {err[worst]:.3e}"
    return got, err


def test_chirp_10k_x_300():
    """The 10k x 300 workload: chirp + N(0, 0.1), LOGSCALES 1-101 Hz, fs=1000, sigma=2."""
    x = chirp_noise(10_000, seed=1)
    sc = O.scales(O.LOGSCALES, 1000, 1, 101, 300)
    _, err = _check(x, sc, 2.0)
    print(f"max rel-L2 {err.max():.2e}")


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 1000, 1023, 1024, 1025, 4097])
def test_lengths(n):
    x = tones_noise(n, seed=n)
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 40)
    _check(x, sc, 2.0)


@pytest.mark.parametrize("fb", [0.7, 1.0, 6.0, 16.0])
def test_bandwidths(fb):
    x = tones_noise(5000, seed=3)
    sc = O.scales(O.LINFREQS, 1000, 1, 400, 64)
    _check(x, sc, fb)


@pytest.mark.parametrize("threads,flags", [(1, 0), (8, 0), (6, 0), (96, 0), (1, 1), (6, 1)])
def test_thread_coverage(threads, flags):
    """Coverage per the reference's thread count and code path, including the
    bins a non-power-of-two count skips (they hold earlier scales' products)."""
    x = tones_noise(20_000, seed=4)
    sc = O.scales(O.LOGSCALES, 1000, 1, 480, 96)
    _check(x, sc, 2.0, threads=threads, flags=flags)


def test_zero_coverage_rows():
    """reference test_lowfrequency setting: fs=100, n=6000, LOGSCALES 0.01-10 Hz,
    fn=300, 8 threads; low rows cover no bins and repeat earlier products."""
    x = tones_noise(6000, fs=100, fmax=31, seed=5)
    sc = O.scales(O.LOGSCALES, 100, 0.01, 10, 300)
    E, _, C = O.rows_meta(8192, sc, 8, True)
    assert (C == 0).sum() > 10
    _check(x, sc, 2.0, threads=8)


def test_unnormalised():
    x = chirp_noise(3000, seed=6)
    sc = O.scales(O.LINSCALES, 1000, 2, 400, 50)
    _check(x, sc, 2.0, normalize=False)


def test_descending_scales():
    x = tones_noise(8000, seed=7)
    sc = O.scales(O.LOGSCALES, 1000, 450, 1, 50)
    _check(x, sc, 2.0)


def test_wide_bands_two_pass():
    """N = 65536 with bands up to N/2: every two-pass band size."""
    x = tones_noise(50_000, seed=8)
    sc = O.scales(O.LOGSCALES, 1000, 0.5, 499, 120)
    _check(x, sc, 2.0)


def test_headline_100k_x_3000_rows():
    """n=100,000 (N=131,072), fn=3000, LOGSCALES 1-450 Hz, fs=1000, sigma=2:
    the full GPU output against the oracle on 40 rows spanning every band size."""
    x = tones_noise(100_000, seed=2024)
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 3000)
    rows = sorted(set(np.linspace(0, 2999, 40).astype(int)) | {0, 1, 628, 629, 2999})
    _check(x, sc, 2.0, rows=rows)
    split = fcwt_amd.Plan(100_000, sc, 2.0).row_split()
    assert split == {"direct": 671, "team": 2329, "two_pass": 0}, split


@pytest.mark.parametrize("teams", [True, False])
@pytest.mark.parametrize("n", [20_000, 50_000, 50_001, 100_000, 150_000, 200_001])
def test_wide_band_paths(n, teams, monkeypatch):
    """N = 2^15 .. 2^18 with bands up to N/2, through the XCD-team kernel
    (at N = 2^18 its 512 x 512 split with the 512-point pass 2) and (with the
    team path switched off) the two-pass fallback, or at N = 2^18 the column
    path; an odd n takes the team kernel's single-value output stores."""
    if not teams:
        monkeypatch.setenv("FCWT_HIP_NO_TEAMS", "1")
    x = tones_noise(n, seed=n)
    sc = O.scales(O.LOGSCALES, 1000, 0.5, 499, 48)
    p = fcwt_amd.Plan(n, sc, 2.0)
    split = p.row_split()
    assert sum(split.values()) == 48
    assert (split["team"] > 0) == teams and (split["two_pass"] > 0) == (not teams), split
    got = run_device(x, sc, 2.0, plan=p)
    ref = O.cwt_fp64(x, sc, 2.0, 1, True, True)
    err = O.rel_l2_rows(got, ref)
    assert err.max() <= TOL, f"row {int(np.argmax(err))}: rel-L2 {err.max():.3e}"


@pytest.mark.parametrize("n", [100_000, 180_000])
@pytest.mark.parametrize("nbuf", [2, 3, 12])
def test_team_buffer_counts(nbuf, n, monkeypatch):
    """The XCD-team kernel with the fewest scratch buffers (2 and 3: pass 2
    one step behind pass 1, no slack) and the most (12), through the
    environment override, at N = 2^17 and 2^18: same rows as the default,
    within the bar."""
    monkeypatch.setenv("FCWT_HIP_TEAM_BUFFERS", str(nbuf))
    x = tones_noise(n, seed=40 + nbuf)
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 120)
    p = fcwt_amd.Plan(x.size, sc, 2.0)
    assert p.row_split()["team"] > 0
    for _ in range(2):
        got = run_device(x, sc, 2.0, plan=p)
        rows = sorted(set(range(0, 120, 7)) | {119})
        ref = O.cwt_fp64(x, sc, 2.0, rows=rows)
        assert O.rel_l2_rows(got[rows], ref).max() <= TOL


def test_direct_rows_inside_team_launch(monkeypatch):
    """With wide rows on the XCD-team kernel, the rows whose band fits one
    work-group (M = 2^6 .. 2^10) run inside that launch as its members' last
    work, with no direct-row launch of their own: bit-identical to the same
    rows through their own launches (FCWT_HIP_NO_FILL=1), and within the bar
    against the oracle for one row of every band size."""
    monkeypatch.setenv("FCWT_HIP_FILL", "1")
    x = tones_noise(100_000, seed=77)
    sc = O.scales(O.LOGSCALES, 1000, 0.2, 450, 200)
    p = fcwt_amd.Plan(x.size, sc, 2.0)
    split = p.row_split()
    assert split["team"] > 0 and split["direct"] > 0
    N = 1 << O.find2power(x.size)
    E, _, _ = O.rows_meta(N, sc, 1, True)
    log2M = np.ceil(np.log2(np.maximum(E, 1))).astype(int)
    assert set(range(6, 11)) <= set(log2M.tolist())  # every band size the launch takes
    p.profile(True)
    got = run_device(x, sc, 2.0, plan=p)
    kinds = p.profile_read()
    p.profile(False)
    assert "row_writer_direct" not in kinds and "row_team" in kinds, kinds
    monkeypatch.delenv("FCWT_HIP_FILL")
    monkeypatch.setenv("FCWT_HIP_NO_FILL", "1")
    q = fcwt_amd.Plan(x.size, sc, 2.0)
    q.profile(True)
    sep = run_device(x, sc, 2.0, plan=q)
    assert "row_writer_direct" in q.profile_read()
    assert np.array_equal(got, sep)
    rows = sorted({int(np.argmax(log2M == m)) for m in set(log2M.tolist())})
    ref = O.cwt_fp64(x, sc, 2.0, rows=rows)
    assert O.rel_l2_rows(got[rows], ref).max() <= TOL
    # the second execute reuses the launch's fill counter
    assert np.array_equal(run_device(x, sc, 2.0, plan=p), got)


def test_team_rows_unaligned_output():
    """An output array that is only 8-byte aligned: the team kernel then
    stores one value at a time instead of 16-byte pairs; same bits as an
    aligned output."""
    import torch

    x = tones_noise(100_000, seed=78)
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 120)
    p = fcwt_amd.Plan(x.size, sc, 2.0)
    assert p.row_split()["team"] > 0
    ref = run_device(x, sc, 2.0, plan=p)
    dx = torch.from_numpy(x).cuda()
    buf = torch.zeros(sc.size * x.size + 1, dtype=torch.complex64, device="cuda")
    p.execute_device(dx.data_ptr(), buf.data_ptr() + 8, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    p.check()
    got = buf[1:].view(sc.size, x.size).cpu().numpy()
    assert np.array_equal(got, ref)
    assert buf[0].item() == 0


@pytest.mark.parametrize("fn", [1, 3, 9])
def test_team_fewer_rows_than_teams(fn):
    """n = 100,000 with 1, 3 or 9 wide rows: teams without rows, and teams
    whose rows hold fewer items than the team has members, so most members
    claim no item and leave at once; every row against the oracle."""
    x = tones_noise(100_000, seed=31 + fn)
    sc = O.scales(O.LINFREQS, 1000, 300, 450, fn)
    p = fcwt_amd.Plan(x.size, sc, 2.0)
    assert p.row_split() == {"direct": 0, "team": fn, "two_pass": 0}
    for _ in range(2):  # the second execute reuses the plan's counters
        got = run_device(x, sc, 2.0, plan=p)
        ref = O.cwt_fp64(x, sc, 2.0)
        assert O.rel_l2_rows(got, ref).max() <= TOL


def test_on_bin_cosine_closed_form():
    """n = N, x = cos(2 pi k0 t/N): row i = (1/N)(N/2) mother[idx_i(k0)] e^{2 pi i k0 t/N}."""
    N, k0 = 16384, 1234
    t = np.arange(N)
    x = np.cos(2 * np.pi * k0 * t / N).astype(np.float32)
    sc = O.scales(O.LOGSCALES, 1000, 20, 450, 64)
    got = run_device(x, sc, 2.0)
    mom = O.mother(2.0, N)
    E, step, C = O.rows_meta(N, sc, 1, True)
    checked = 0
    for i in range(len(sc)):
        if k0 < C[i]:
            idx = int(min(np.float32(N - 1), step[i] * np.float32(k0)))
            expect = 0.5 * float(mom[idx]) * np.exp(2j * np.pi * k0 * t / N)
            if np.abs(expect).max() > 1e-3:
                assert np.abs(got[i] - expect).max() <= 1e-5 * np.abs(mom).max(), i
                checked += 1
    assert checked > 5


def test_linearity_full_size():
    """cwt(a + b) = cwt(a) + cwt(b) at the headline size (float32 tolerance)."""
    a = tones_noise(100_000, seed=1)
    b = chirp_noise(100_000, seed=2)
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 3000)
    p = fcwt_amd.Plan(100_000, sc, 2.0)
    ya = run_device(a, sc, 2.0, plan=p)
    yb = run_device(b, sc, 2.0, plan=p)
    yab = run_device(a + b, sc, 2.0, plan=p)
    err = O.rel_l2_rows(yab, ya + yb)
    assert err.max() < 3e-5


def test_host_api_matches_device_api():
    x = chirp_noise(10_000, seed=9)
    morl = fcwt_amd.Morlet(2.0)
    scs = fcwt_amd.Scales(morl, fcwt_amd.FCWT_LOGSCALES, 1000, 1, 101, 300)
    fc = fcwt_amd.FCWT(morl, 1, False, True)
    out = np.zeros((300, x.size), np.complex64)
    fc.cwt(x, scs, out)
    dev = run_device(x, scs.scales, 2.0)
    assert np.array_equal(out, dev)
    freqs, out2 = fcwt_amd.cwt(x, 1000, 1, 101, 300, scaling="log")
    assert np.array_equal(out2, dev)


@pytest.mark.parametrize("n", [1000, 8192, 16384, 100_000, 1 << 22])
@pytest.mark.parametrize("cplx", [False, True])
def test_forward_spectrum_accuracy(n, cplx):
    """The forward transform alone (fcwt_hip_spectrum_device) against the
    float64 DFT: relative L2 well under 1e-6 at every size and path (one
    work-group radix-16 FFT up to 2^14 points, four-step beyond).  Rows that
    only read bins far below the spectrum's peak inherit this absolute error,
    which is why a few low-frequency rows of test_zero_coverage_rows sit near
    the 1e-5 bar (a float32 FFT of the reference would do the same)."""
    import torch

    x = tones_noise(n, fs=1000, fmax=310, seed=n)
    x = x.astype(np.complex64) * np.complex64(1 - 0.5j) if cplx else x
    p = fcwt_amd.Plan(n, np.array([2.0], np.float32), 2.0)
    dx = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    sp = torch.empty(p.N // 2, dtype=torch.complex64, device="cuda")
    p.spectrum_device(dx.data_ptr(), sp.data_ptr(), torch.cuda.current_stream().cuda_stream, cplx)
    torch.cuda.synchronize()
    got = sp.cpu().numpy()
    ref = O.forward_fp64(x, p.N)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= 5e-7, err
