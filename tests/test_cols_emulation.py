"""CPU emulation of the column path's index arithmetic. Pass 1
(fcwt_amd/csrc/rows_big.h: cols_skip, cols_fft, cols_pass): for every F of
the path and bands A <= F/2, the passes that only replicate are skipped, the
first pass run reads position p of a column as z[p >> log2S], the passes in
between are radix-2^(log2F mod 4) then radix-16 Stockham steps with the
kernel's twiddle exponents, and the last pass writes position
dst + k Ns.  The result must be the inverse DFT of the band,
sum_{q < A} z[q] e^{+2 pi i q f / F}, to float64 rounding.  Pass 2 at
B = 512 (cols_pass2_b512): the two register stages, the LDS exchange
indices and the lane-pair radix-2 step, against the direct sum over qb.
No GPU."""
import numpy as np
import pytest


def cols_skip(log2F, A):
    first = log2F % 4
    npass = (first > 0) + (log2F - first) // 4
    skip, l2, step = 0, 0, first if first > 0 else 4
    while skip < npass and (A << (l2 + step)) <= (1 << log2F):
        l2 += step
        skip += 1
        step = 4
    return skip, l2, npass


def passes(log2F):
    first = log2F % 4
    npass = (first > 0) + (log2F - first) // 4
    out = []
    for j in range(npass):
        log2R = first if (j == 0 and first > 0) else 4
        log2NS = 0 if j == 0 else (first + 4 * (j - 1) if first > 0 else 4 * j)
        out.append((log2R, log2NS))
    return out


def emulate_column(log2F, z):
    F = 1 << log2F
    A = z.size
    skip, log2S, npass = cols_skip(log2F, A)
    p = np.arange(F)
    gq = p >> log2S
    x = np.where(gq < A, np.concatenate([z, np.zeros(F, complex)])[np.minimum(gq, F - 1)], 0)
    if skip >= npass:  # every pass replicates
        return x
    for log2R, log2NS in passes(log2F)[skip:]:
        R, Ns = 1 << log2R, 1 << log2NS
        log2GPC = log2F - log2R
        gl = np.arange(1 << log2GPC)
        k = np.arange(R)
        v = x[gl[:, None] + (k[None, :] << log2GPC)]
        if log2NS > 0:
            jj = gl & (Ns - 1)
            sh = log2F - log2NS - log2R
            v = v * np.exp(2j * np.pi * (((jj[:, None] * k[None, :]) << sh) % F) / F)
        v = np.fft.ifft(v, axis=1) * R  # sum_j v[j] e^{+2 pi i jk / R}
        dst = ((gl >> log2NS) << (log2NS + log2R)) + (gl & (Ns - 1))
        y = np.empty(F, complex)
        y[dst[:, None] + (k[None, :] << log2NS)] = v
        x = y
    return x


@pytest.mark.parametrize("log2F", range(5, 15))
def test_pass1_transform_every_F(log2F):
    F = 1 << log2F
    rng = np.random.default_rng(log2F)
    for A in sorted({1, 2, 3, 4, 8, 16, 17, 32, 100, F // 32, F // 16 + 1, F // 4, F // 2} - {0}):
        if A > F // 2:
            continue
        z = rng.standard_normal(A) + 1j * rng.standard_normal(A)
        got = emulate_column(log2F, z)
        ref = np.fft.ifft(np.concatenate([z, np.zeros(F - A)])) * F
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 1e-12, (log2F, A, err)


def test_pass2_512_point_decomposition():
    """cols_pass2_b512 (rows_big.h): 512 threads per 16 f, qb = j + 32 k,
    m = m1 + 16 m2.  Stage 1 (thread g = tid & 15, j = tid >> 4): W_N^{32 k f}
    from the [k][g] table, a 16-point DFT over k, then c W_N^{j f} W_512^{j m1}
    (W_512^{j m1} = W_N^{(j m1) << log2F}, the [m1][j] table); the exchange
    xb[m1][g][j]; stage 2 (g, e = (tid >> 4) & 1, m1 = tid >> 5): a 16-point
    DFT over j' of xb[m1][g][2 j' + e]; then lane e = 1 multiplies by
    W_32^{m2'} (the table entry [m2'][16]) and the lanes 16 apart swap:
    e = 0 writes E_0 + W E_1 at m = m1 + 16 m2', e = 1 writes E_0 - W E_1 at
    m + 256.  Against y[f + F m] = sum_qb W_512^{qb m} W_N^{qb f} U[qb][f]."""
    rng = np.random.default_rng(5)
    log2F = 6
    F, B = 1 << log2F, 512
    N = B * F
    U = rng.standard_normal((B, F)) + 1j * rng.standard_normal((B, F))

    def W(X, r):
        return np.exp(2j * np.pi * np.asarray(r) / X)

    ref = np.zeros(N, complex)
    for f in range(F):
        ref[f + F * np.arange(B)] = np.fft.ifft(U[:, f] * W(N, np.arange(B) * f)) * B
    wb = W(N, (np.arange(16)[:, None] * np.arange(32)[None, :]) << log2F)  # [m1][j]
    got = np.zeros(N, complex)
    k = np.arange(16)
    for f0 in range(0, F, 16):
        xb = np.zeros((16, 16, 33), complex)
        for tid in range(512):
            g, j = tid & 15, tid >> 4
            f = f0 + g
            v = U[j + 32 * k, f] * W(N, (32 * k * f) % N)
            v = np.fft.ifft(v) * 16
            v = v * W(N, j * f) * wb[:, j]
            xb[:, g, j] = v
        E = np.zeros((512, 16), complex)
        for tid in range(512):
            g, e, m1 = tid & 15, (tid >> 4) & 1, tid >> 5
            E[tid] = np.fft.ifft(xb[m1, g, 2 * np.arange(16) + e]) * 16
        x = E.copy()
        odd = ((np.arange(512) >> 4) & 1) == 1
        x[odd] *= wb[np.arange(16), 16][None, :]
        for tid in range(512):
            g, e, m1 = tid & 15, (tid >> 4) & 1, tid >> 5
            pr = x[tid ^ 16]
            o = pr - x[tid] if e else x[tid] + pr
            got[f0 + g + F * (m1 + 16 * np.arange(16) + 256 * e)] = o
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12


def test_team_pass2_512_point_decomposition():
    """team_pass2_b512 (rows_team.h): 256 threads per 8 f of a team's scratch
    U[f][qb] (row stride 512), qb = j + 32 k, m = m1 + 16 m2.  Stage 1 (thread
    g = tid >> 5, j = tid & 31): W_N^{32 k f} from the [g][k] table, a 16-point
    DFT over k, then c W_N^{j f} W_512^{j m1}; the exchange xb[g][m1][j] at
    g * 548 + m1 * 34 + j; stage 2 (g = tid & 7, e = (tid >> 4) & 1,
    m1 = ((tid >> 3) & 1) + 2 (tid >> 5)): a 16-point DFT over j' of
    xb[g][m1][2 j' + e]; lane e = 1 multiplies by W_32^{m2'}, the lanes 16
    apart (tid ^ 16: same g and m1, the other e) swap, e = 0 writes
    E_0 + W E_1 at m = m1 + 16 m2' and e = 1 writes E_0 - W E_1 at m + 256.
    Every exchange index stays inside the 4768-value LDS buffer.  Against
    y[f + F m] = sum_qb W_512^{qb m} W_N^{qb f} U[f][qb]."""
    rng = np.random.default_rng(7)
    log2F = 5
    F, B = 1 << log2F, 512
    N = B * F
    U = rng.standard_normal((F, B)) + 1j * rng.standard_normal((F, B))

    def W(X, r):
        return np.exp(2j * np.pi * np.asarray(r) / X)

    ref = np.zeros(N, complex)
    for f in range(F):
        ref[f + F * np.arange(B)] = np.fft.ifft(U[f] * W(N, np.arange(B) * f)) * B
    wb = W(N, (np.arange(16)[:, None] * np.arange(32)[None, :]) << log2F)  # [m1][j]
    JS, XS = 34, 16 * 34 + 4
    got = np.full(N, np.nan, complex)
    k = np.arange(16)
    tids = np.arange(256)
    assert (7 * XS + 15 * JS + 31) < 4768
    for f0 in range(0, F, 8):
        xb = np.full(4768, np.nan, complex)
        for tid in tids:
            g, j = tid >> 5, tid & 31
            f = f0 + g
            v = U[f, j + 32 * k] * W(N, (32 * k * f) % N)
            v = np.fft.ifft(v) * 16
            v = v * W(N, j * f) * wb[:, j]
            xb[g * XS + k * JS + j] = v  # k here runs over m1
        E = np.zeros((256, 16), complex)
        for tid in tids:
            g, e, m1 = tid & 7, (tid >> 4) & 1, ((tid >> 3) & 1) + 2 * (tid >> 5)
            E[tid] = np.fft.ifft(xb[g * XS + m1 * JS + 2 * np.arange(16) + e]) * 16
        x = E.copy()
        odd = ((tids >> 4) & 1) == 1
        x[odd] *= wb[np.arange(16), 16][None, :]
        seen = set()
        for tid in tids:
            g, e, m1 = tid & 7, (tid >> 4) & 1, ((tid >> 3) & 1) + 2 * (tid >> 5)
            p = tid ^ 16
            assert (p & 7, ((p >> 3) & 1) + 2 * (p >> 5)) == (g, m1) and ((p >> 4) & 1) != e
            pr = x[p]
            o = pr - x[tid] if e else x[tid] + pr
            t = f0 + g + F * (m1 + 16 * np.arange(16) + 256 * e)
            seen.update(t.tolist())
            got[t] = o
        assert len(seen) == 8 * 512
    assert not np.isnan(got).any()
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
