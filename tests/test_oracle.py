"""Pins the oracle (oracle/) before anything is checked against it.

* FP32 host arithmetic (scales, frequencies, mother table, per-scale coverage,
  window multiply with O1 reuse, time-domain wavelet) against vectors the
  reference's own fcwt.cpp produced (tests/golden/ref_vectors.npz, written by
  tests/golden/make_golden.py), and live against the reference probes when
  they are built here (oracle/_ref).
* Float64 transforms against a brute-force DFT, the on-bin cosine closed form,
  and DFT identities (linearity, Parseval, impulse, circular shift).
* Planted mistakes in the oracle (a dropped window term, a wrong index, a sign
  flip) must fail these checks.
"""
import hashlib
import pathlib
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

GOLD = np.load(pathlib.Path(__file__).parent / "golden" / "ref_vectors.npz")
REF = pathlib.Path(O.HERE) / "_ref"


def _cases(prefix):
    k = 0
    while f"{prefix}_{k}_args" in GOLD:
        yield k, GOLD[f"{prefix}_{k}_args"]
        k += 1


# ----------------------------------------------------------- golden vectors
def test_scales_match_reference_vectors():
    n = 0
    for k, args in _cases("scales"):
        t, fs, f0, f1, fn = args
        s = O.scales(int(t), int(fs), float(f0), float(f1), int(fn))
        assert np.array_equal(s.view(np.uint32), GOLD[f"scales_{k}_scales"].view(np.uint32)), k
        fr = O.frequencies(s, int(fs))
        assert np.array_equal(fr.view(np.uint32), GOLD[f"scales_{k}_freqs"].view(np.uint32)), k
        n += 1
    assert n >= 10


def test_mother_matches_reference_vectors():
    for k, (fb, N) in _cases("mother"):
        m = O.mother(float(fb), int(N))
        assert hashlib.sha256(m.tobytes()).digest() == GOLD[f"mother_{k}_sha256"].tobytes(), (fb, N)
        if f"mother_{k}_values" in GOLD:
            assert np.array_equal(m, GOLD[f"mother_{k}_values"])


def test_window_multiply_matches_reference_vectors():
    N = int(GOLD["window_N"])
    sc = GOLD["window_scales"]
    ihat = GOLD["window_ihat"]
    seen_zero_cov = False
    for k, (fb, avx, T) in _cases("window"):
        ref = GOLD[f"window_{k}_o1"]
        mom = O.mother(float(fb), N)
        o1 = np.zeros(N, np.complex64)
        for i, s in enumerate(sc):
            O.window_multiply(ihat, o1, mom, float(s), N, int(T), bool(avx))
            assert np.array_equal(o1.view(np.uint64), ref[i].view(np.uint64)), (fb, avx, T, i)
        E, _, C = O.rows_meta(N, sc, int(T), bool(avx))
        seen_zero_cov |= bool((C == 0).any())
    assert seen_zero_cov  # the fixture exercises scales the reference covers with no bins


def test_time_wavelet_matches_reference_vectors():
    for k, (fb, scale, pn) in _cases("wavelet"):
        w = O.time_wavelet(float(fb), float(scale), int(pn))
        assert np.array_equal(w.view(np.uint64), GOLD[f"wavelet_{k}_values"].view(np.uint64)), k
        width = int(np.float32(fb) * np.float32(scale) * np.float32(3.0))
        assert width == int(GOLD[f"wavelet_{k}_support"])


# --------------------------------------------- live reference probe (here only)
needs_ref = pytest.mark.skipif(not (REF / "ref_probe_avx").exists(), reason="reference probe not built")


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_window_multiply_matches_reference_probe_random(seed):
    rng = np.random.default_rng(seed)
    N = int(2 ** rng.integers(6, 13))
    fn = int(rng.integers(3, 40))
    fb = float(np.float32(rng.uniform(0.3, 20)))
    T = int(rng.choice([1, 2, 3, 5, 8, 12, 64]))
    avx = bool(rng.integers(0, 2))
    lo, hi = sorted(rng.uniform(0.5, N * 1.5, size=2))
    sc = np.sort(rng.uniform(lo, hi + 1, size=fn)).astype(np.float32)
    if rng.integers(0, 2):
        sc = sc[::-1].copy()
    ihat = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    binary = REF / ("ref_probe_avx" if avx else "ref_probe_scalar")
    raw = subprocess.run([str(binary), "window", repr(fb), str(N), str(T), str(fn)],
                         input=sc.tobytes() + ihat.tobytes(), capture_output=True, check=True).stdout
    ref = np.frombuffer(raw, np.complex64).reshape(fn, N)
    mom = O.mother(fb, N)
    o1 = np.zeros(N, np.complex64)
    for i, s in enumerate(sc):
        O.window_multiply(ihat, o1, mom, float(s), N, T, avx)
        assert np.array_equal(o1.view(np.uint64), ref[i].view(np.uint64)), i


@needs_ref
@pytest.mark.parametrize("args", [(1, 44100, 20.0, 20000.0, 512), (2, 250, 0.5, 120.0, 999), (0, 8000, 3.0, 3999.0, 33)])
def test_scales_match_reference_probe(args):
    t, fs, f0, f1, fn = args
    raw = subprocess.run([str(REF / "ref_probe_avx"), "scales", str(t), str(fs), repr(f0), repr(f1), str(fn)],
                         capture_output=True, check=True).stdout
    ref = np.frombuffer(raw, np.float32)
    s = O.scales(t, fs, f0, f1, fn)
    assert np.array_equal(s, ref[:fn]) and np.array_equal(O.frequencies(s, fs), ref[fn:])


# ------------------------------------------------------- closed forms/invariants
def test_find2power():
    for n in [1, 2, 3, 4, 5, 1000, 1023, 1024, 1025, 100000, 4194304]:
        m = O.find2power(n)
        assert 2 ** m >= n and (m == 0 or 2 ** (m - 1) < n)


def test_mother_closed_form():
    for fb, N in [(2.0, 4096), (6.0, 1024), (0.5, 256)]:
        m = O.mother(fb, N)
        peak = np.float32(np.sqrt(np.float32(2 * np.float32(np.pi)))) * np.float32(0.75112554446)
        assert m[N // 2] == pytest.approx(float(peak), rel=1e-6)
        assert int(np.argmax(m)) == N // 2
        w = np.arange(N)
        expect = float(peak) * np.exp(-0.5 * (2 * np.pi * fb * (2 * w / N - 1)) ** 2)
        np.testing.assert_allclose(m, expect, rtol=2e-4 * max(1.0, fb), atol=1e-30)


def test_scales_invariants():
    s = O.scales(O.LOGSCALES, 1000, 1, 450, 3000)
    assert np.all(np.diff(s) > 0)
    f = O.frequencies(s, 1000)
    assert f[0] == pytest.approx(450, rel=1e-6) and f[-1] == pytest.approx(1, rel=1e-5)
    s = O.scales(O.LINFREQS, 1000, 1, 20, 1000)
    f = O.frequencies(s, 1000)
    # top frequency is f1 - (f1-f0)/fn, not f1 (fcwt.cpp:170-171)
    assert f[-1] == pytest.approx(1.0, rel=1e-6) and f[0] == pytest.approx(20 - 19 / 1000, rel=1e-5)
    assert np.allclose(np.diff(f), -19 / 1000, rtol=1e-3)


def test_coverage_closed_form():
    N = 131072
    s = O.scales(O.LOGSCALES, 1000, 1, 450, 3000)
    E, step, C = O.rows_meta(N, s, 1, True)
    Eexp = np.minimum(N / 2, 2 * N / s.astype(np.float64))
    assert np.all(np.abs(E - Eexp) <= 1)
    assert np.all(C == 4 * (E // 4))
    _, _, Cs = O.rows_meta(N, s, 1, False)
    assert np.all(Cs == E)
    # non-power-of-two thread counts skip near-Nyquist bins of capped rows
    _, _, C96 = O.rows_meta(N, s, 96, True)
    assert C96[0] < N // 2 and np.all(C96 <= C)


# -------------------------------------------------------------- transforms
def _dft(x, N, sign):
    t = np.arange(len(x))
    k = np.arange(N)
    return (np.asarray(x, np.complex128)[None, :] * np.exp(sign * 2j * np.pi * np.outer(k, t) / N)).sum(axis=1)


@pytest.mark.parametrize("n", [1, 2, 7, 64, 100, 255])
def test_forward_fp64_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n).astype(np.float32)
    N = 1 << O.find2power(n)
    X = O.forward_fp64(x, N)
    np.testing.assert_allclose(X, _dft(x, N, -1)[: N // 2], rtol=0, atol=1e-10 * max(1, np.abs(x).sum()))


def test_forward_identities():
    rng = np.random.default_rng(3)
    N = 1024
    a, b = rng.integers(-1000, 1000, size=(2, 1000)).astype(np.float32)  # a + b exact in float32
    # linearity
    np.testing.assert_allclose(O.forward_fp64(a + b, N), O.forward_fp64(a, N) + O.forward_fp64(b, N), atol=1e-9)
    # impulse at t0 -> e^{-2 pi i k t0 / N}
    x = np.zeros(1000, np.float32)
    x[37] = 1
    k = np.arange(N // 2)
    np.testing.assert_allclose(O.forward_fp64(x, N), np.exp(-2j * np.pi * k * 37 / N), atol=1e-12)


@pytest.mark.parametrize("n", [1, 3, 64, 100, 255])
def test_forward_fp64_complex_brute_force(n):
    """Complex input: the c2c transform of the zero-padded signal (fcwt.cpp:486-490)."""
    rng = np.random.default_rng(100 + n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    N = 1 << O.find2power(n)
    X = O.forward_fp64(x, N)
    np.testing.assert_allclose(X, _dft(x, N, -1)[: N // 2], rtol=0, atol=1e-10 * max(1, np.abs(x).sum()))


def test_cwt_fp64_complex_closed_forms():
    """n = N.  x = e^{+2 pi i k0 t/N}: X[k0] = N, so row i is c N mother[idx_i(k0)]
    e^{2 pi i k0 t/N} when k0 < C_i.  x = e^{-2 pi i k0 t/N} sits at bin N - k0
    >= N/2, which no row reads (Morlet is one-sided): every row is zero.
    A real signal given as complex gives the real-input rows."""
    N, k0 = 2048, 200
    t = np.arange(N)
    sc = O.scales(O.LOGSCALES, 1000, 5, 450, 30)
    mom = O.mother(2.0, N)
    E, step, C = O.rows_meta(N, sc, 1, True)
    pos = np.exp(2j * np.pi * k0 * t / N).astype(np.complex64)
    got = O.cwt_fp64(pos, sc, 2.0)
    checked = 0
    for i in range(len(sc)):
        if k0 < C[i]:
            idx = int(min(np.float32(N - 1), step[i] * np.float32(k0)))
            expect = float(mom[idx]) * pos.astype(np.complex128)  # c N = 1
            assert np.abs(got[i] - expect).max() <= 1e-6 * float(mom.max()), i
            checked += 1
    assert checked > 3
    neg = np.exp(-2j * np.pi * k0 * t / N).astype(np.complex64)
    assert np.abs(O.cwt_fp64(neg, sc, 2.0)).max() <= 1e-6 * float(mom.max())
    xr = np.random.default_rng(7).standard_normal(N - 5).astype(np.float32)
    np.testing.assert_allclose(O.cwt_fp64(xr.astype(np.complex64), sc, 2.0), O.cwt_fp64(xr, sc, 2.0),
                               rtol=0, atol=1e-12 * np.abs(xr).sum())


def _tiny_problem(seed, n, fn, fb, lowcov=False):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n).astype(np.float32)
    fs = 100
    f0 = 0.05 if lowcov else 1.0
    sc = O.scales(O.LOGSCALES, fs, f0, 45.0, fn)
    return x, sc


@pytest.mark.parametrize("seed,n,fn,fb,T,avx,norm,lowcov", [
    (0, 64, 12, 2.0, 1, True, True, False),
    (1, 100, 9, 0.7, 8, True, False, False),
    (2, 257, 7, 6.0, 6, True, True, False),
    (3, 200, 16, 2.0, 1, False, True, True),
    (4, 300, 20, 0.6, 3, True, True, True),
])
def test_cwt_fp64_brute_force(seed, n, fn, fb, T, avx, norm, lowcov):
    x, sc = _tiny_problem(seed, n, fn, fb, lowcov)
    got = O.cwt_fp64(x, sc, fb, T, avx, norm)
    ref = O.brute_force_cwt(x, sc, fb, T, avx, norm)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * scale)
    # a subset of rows gives the same rows (stale bins are tracked across all rows)
    sub = O.cwt_fp64(x, sc, fb, T, avx, norm, rows=[fn - 1, 2])
    np.testing.assert_allclose(sub, got[[2, fn - 1]], rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("seed,n,fn,fb,T", [(10, 64, 10, 2.0, 1), (11, 150, 8, 1.0, 8), (12, 257, 6, 6.0, 6)])
def test_cwt_fp64_complex_brute_force(seed, n, fn, fb, T):
    x, sc = _tiny_problem(seed, n, fn, fb)
    x = (x + 1j * np.random.default_rng(seed).standard_normal(n)).astype(np.complex64)
    got = O.cwt_fp64(x, sc, fb, T)
    ref = O.brute_force_cwt(x, sc, fb, T)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_cwt_on_bin_cosine_closed_form():
    # n = N, x = cos(2 pi k0 t / N): X[k0] = N/2 exactly and
    # W_i[t] = c * (N/2) * mother[idx_i(k0)] * e^{2 pi i k0 t / N}  when k0 < C_i.
    N = 4096
    k0 = 300
    t = np.arange(N)
    x = np.cos(2 * np.pi * k0 * t / N).astype(np.float32)
    sc = O.scales(O.LOGSCALES, 1000, 5, 450, 40)
    got = O.cwt_fp64(x, sc, 2.0, 1, True, True)
    mom = O.mother(2.0, N)
    E, step, C = O.rows_meta(N, sc, 1, True)
    for i in range(len(sc)):
        if k0 < C[i]:
            idx = int(min(np.float32(N - 1), step[i] * np.float32(k0)))
            expect = (1 / N) * (N / 2) * float(mom[idx]) * np.exp(2j * np.pi * k0 * t / N)
            err = np.abs(got[i] - expect).max()
            # float32 input rounding makes the other bins ~1e-7 of the peak
            assert err <= 1e-6 * (N / 2) / N * float(mom.max()) + 1e-6 * np.abs(expect).max(), i


def test_cwt_fp32_port_matches_fp64():
    x, sc = _tiny_problem(5, 3000, 30, 2.0)
    a = O.cwt_fp32_port(x, sc, 2.0)
    b = O.cwt_fp64(x, sc, 2.0)
    assert O.rel_l2_rows(a, b).max() < 1e-5
    # blocks of rows on several threads give the same values (the O1 reuse
    # crosses block edges: block = 7 splits the 30 rows unevenly)
    c = O.cwt_fp32_port(x, sc, 2.0, workers=4, block=7)
    assert np.array_equal(a, c)


def _o1_composition_matches(case):
    """cwt_fp64's spectra (o1_rows) rounded to float32 equal the reference's O1
    snapshots bit for bit: X and the window are float32, so their float64
    product is exact and rounds to the reference's float32 product."""
    fb, avx, T = GOLD[f"window_{case}_args"]
    N = int(GOLD["window_N"])
    ihat = GOLD["window_ihat"]
    ref = GOLD[f"window_{case}_o1"]
    mom = O.mother(float(fb), N)
    for i, O1 in O.o1_rows(ihat.astype(np.complex128), GOLD["window_scales"], mom, N, int(T), bool(avx)):
        if not np.array_equal(O1.astype(np.complex64).view(np.uint64), ref[i, : N // 2].view(np.uint64)):
            return False
        if np.any(ref[i, N // 2:] != 0):
            return False
    return True


def test_o1_composition_matches_reference_vectors():
    for k, _ in _cases("window"):
        assert _o1_composition_matches(k), k


def _brute_force_matches(fb=2.0, norm=True):
    x, sc = _tiny_problem(7, 128, 10, fb)
    ref = O.brute_force_cwt(x, sc, fb, normalize=norm)
    got = O.cwt_fp64(x, sc, fb, normalize=norm)
    return np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


# --------------------------------------------------------- planted mistakes
def test_planted_mistakes_are_caught(monkeypatch):
    """Plausible oracle bugs (a shifted window index, a wrong DFT sign, a
    dropped normalisation) each fail one of the checks above."""
    assert _o1_composition_matches(0) and _brute_force_matches()
    real_row_window = O.row_window

    def shifted_window(mom, scale, N, threads=1, avx=True):
        w, C = real_row_window(mom, scale, N, threads, avx)
        return np.roll(w, 1), C

    with monkeypatch.context() as m:
        m.setattr(O, "row_window", shifted_window)
        assert not _o1_composition_matches(0)

    def wrong_sign_forward(x, N):
        xp = np.zeros(N)
        xp[: x.size] = x
        return np.fft.fft(xp)[: N // 2].conj()

    with monkeypatch.context() as m:
        m.setattr(O, "forward_fp64", wrong_sign_forward)
        assert not _brute_force_matches()

    real_ifft = np.fft.ifft
    with monkeypatch.context() as m:
        m.setattr(np.fft, "ifft", lambda a: real_ifft(a) * a.size)  # normalisation dropped
        assert not _brute_force_matches()


# ------------------------------------------------------------------ adjoint
def _cwt_matrix(n, sc, fb, complex_input):
    """The forward as an explicit (fn*n) x n matrix, from unit inputs."""
    cols = []
    for t in range(n):
        e = np.zeros(n, np.complex64 if complex_input else np.float32)
        e[t] = 1
        cols.append(O.cwt_fp64(e, sc, fb).reshape(-1))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("complex_input", [False, True])
def test_adjoint_is_the_transpose(complex_input):
    """Brute force on a tiny case: xbar = (Re of) A^H G with A the explicit
    forward matrix; plus the dot-product identity Re<W(x), G> = Re<x, xbar>."""
    rng = np.random.default_rng(21)
    n, fn, fb = 50, 7, 2.0
    sc = O.scales(O.LOGSCALES, 100, 0.5, 45, fn)  # includes zero-coverage (stale) rows
    A = _cwt_matrix(n, sc, fb, complex_input)
    G = (rng.standard_normal((fn, n)) + 1j * rng.standard_normal((fn, n)))
    ref = A.conj().T @ G.reshape(-1)
    got = O.adjoint_fp64(G, sc, fb, complex_input=complex_input)
    np.testing.assert_allclose(got, ref if complex_input else ref.real, rtol=0, atol=1e-12 * np.abs(ref).max())
    x = rng.standard_normal(n) + (1j * rng.standard_normal(n) if complex_input else 0)
    W = O.cwt_fp64(x.astype(np.complex64 if complex_input else np.float32), sc, fb)
    x32 = x.astype(np.complex64 if complex_input else np.float32).astype(np.complex128)
    lhs = np.real(np.vdot(W.reshape(-1), G.reshape(-1)))
    rhs = np.real(np.vdot(x32, got))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_adjoint_matches_finite_differences():
    """L = sum w |W|^2: central differences of L (float64) at a few samples
    equal the adjoint of G = 2 w W."""
    rng = np.random.default_rng(5)
    n, fn, fb = 300, 12, 2.0
    sc = O.scales(O.LOGSCALES, 1000, 10, 450, fn)
    x = rng.standard_normal(n)
    wts = rng.uniform(0, 1, (fn, n))

    # cwt_fp64 rounds its input to float32; differentiate the float64 map instead
    def W64(v):
        N = 1 << O.find2power(n)
        X = np.fft.rfft(np.pad(v, (0, N - n)))[: N // 2]
        rows = []
        for i, w in O.effective_windows(sc, fb, N):
            full = np.zeros(N, np.complex128)
            full[: N // 2] = X * w
            rows.append(np.fft.ifft(full)[:n])
        return np.stack(rows)

    L = lambda v: float((wts * np.abs(W64(v)) ** 2).sum())
    G = 2 * wts * W64(x)
    xbar = O.adjoint_fp64(G, sc, fb)
    h = 1e-3
    for t in rng.choice(n, 8, replace=False):
        e = np.zeros(n)
        e[t] = h
        fd = (L(x + e) - L(x - e)) / (2 * h)
        assert abs(fd - xbar[t]) <= 1e-6 * max(1.0, abs(fd)), t
