"""CPU emulation of the direct kernel's exchange at M = 256, 512 and 1024
(fcwt_amd/csrc/rows_direct.h, direct_block): the LDS exchange indices of
the two rounds over the halves of k1, the thread mapping of the second-stage
DFTs (g in the low lane bits) and, at M = 1024, the 16-point halves on lanes
16 apart with the radix-2 step between them, for the separate kernel (16
phases) and the team launch's blocks (8 phases).  Every output index is
written once and the block's outputs equal the direct sum
y[l + L m] = sum_q Z[q] W_N^{q (l + L m)}, to float64 rounding.  No GPU."""
import numpy as np
import pytest


def W(X, r):
    return np.exp(2j * np.pi * np.asarray(r, dtype=np.float64) / X)


def stage1(Z, M, N, G, l0):
    """v[g][j][k1] after the R1 = 32-point DFT over k and the factors
    W_N^{j l} W_M^{j k1} (the kernel's pre-twiddles, DFT, base and wm)."""
    R1, P2 = 32, M // 32
    v = np.zeros((G, P2, R1), complex)
    k = np.arange(R1)
    for g in range(G):
        l = l0 + g
        for j in range(P2):
            x = Z[j + P2 * k] * W(N, (k * l * P2) % N)
            d = np.array([np.sum(x * W(R1, k * k1)) for k1 in range(R1)])
            v[g, j] = d * W(N, (j * l) % N) * W(M, j * np.arange(R1))
    return v


@pytest.mark.parametrize("M,G", [(256, 32), (512, 16), (1024, 16), (1024, 8)])
def test_direct_exchange_and_register_stores(M, G):
    rng = np.random.default_rng(M + G)
    N = 1 << 14
    L = N // M
    l0 = 3 * G
    Z = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    v = stage1(Z, M, N, G, l0)
    P2 = M // 32
    NT = G * P2
    out = {}
    if M == 256:
        GS, KS = 9, G * 9
        assert 16 * KS <= 4768
        for h in range(2):
            buf = np.full(16 * KS, np.nan, complex)
            for tid in range(NT):  # writers (g, j)
                g, j = tid // P2, tid % P2
                for kk in range(16):
                    buf[kk * KS + g * GS + j] = v[g, j, 16 * h + kk]
            for tid in range(NT):  # readers (g2, gb, kr): k1' = kr, kr + 8 of phase 16 gb + g2
                g2, gb, kr = tid & 15, (tid >> 4) & 1, tid >> 5
                gp = 16 * gb + g2
                for d in range(2):
                    kk = kr + 8 * d
                    w = buf[kk * KS + gp * GS + np.arange(8)]
                    y = np.array([np.sum(w * W(8, np.arange(8) * k2)) for k2 in range(8)])
                    for k2 in range(8):
                        t = l0 + gp + L * (16 * h + kk + 32 * k2)
                        assert t not in out
                        out[t] = y[k2]
    elif M == 512:
        KS = G * 17
        for h in range(2):
            buf = np.full(4352, np.nan, complex)
            for tid in range(NT):  # writers (g, j)
                g, j = tid // P2, tid % P2
                for kk in range(16):
                    buf[kk * KS + g * 17 + j] = v[g, j, 16 * h + kk]
            for tid in range(NT):  # readers (g2, kr)
                g2, kr = tid & 15, tid >> 4
                w = buf[kr * KS + g2 * 17 + np.arange(16)]
                y = np.array([np.sum(w * W(16, np.arange(16) * k2)) for k2 in range(16)])
                for k2 in range(16):
                    t = l0 + g2 + L * (16 * h + kr + 32 * k2)
                    assert t not in out
                    out[t] = y[k2]
    else:
        JS, KS = 33, G * 33
        assert 16 * KS + 16 <= (8704 if G == 16 else 4608)
        w32 = W(N, 32 * np.arange(16) * L)
        for h in range(2):
            buf = np.full(16 * KS, np.nan, complex)
            for tid in range(NT):
                g, j = tid // P2, tid % P2
                for kk in range(16):
                    buf[kk * KS + g * JS + j] = v[g, j, 16 * h + kk]
            E = {}
            for tid in range(NT):
                g2, e = tid & (G - 1), (tid >> 4) & 1
                kr = tid >> 5 if G == 16 else ((tid >> 3) & 1) + 2 * (tid >> 5)
                w = buf[kr * KS + g2 * JS + 2 * np.arange(16) + e]
                E[tid] = (g2, e, kr, np.array([np.sum(w * W(16, np.arange(16) * m2)) for m2 in range(16)]))
            for tid in range(NT):
                g2, e, kr, x = E[tid]
                p = tid ^ 16
                pg2, pe, pkr, px = E[p]
                assert (pg2, pkr) == (g2, kr) and pe != e
                xs = x * np.where(np.arange(16) > 0, w32, 1) if e else x
                ps = px * np.where(np.arange(16) > 0, w32, 1) if pe else px
                o = ps - xs if e else xs + ps
                for m2 in range(16):
                    t = l0 + g2 + L * (16 * h + kr + 32 * (m2 + 16 * e))
                    assert t not in out
                    out[t] = o[m2]
    assert len(out) == G * M
    q = np.arange(M)
    for t, val in out.items():
        ref = np.sum(Z * W(N, (q * t) % N))
        assert abs(val - ref) <= 1e-9 * np.abs(Z).sum(), (t, val, ref)
