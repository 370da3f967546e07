"""GPU parity on seeded random plans: signal length (log-uniform in 64 ..
600,000; the drawn cases pad to 2^9 ... 2^19 and reach the direct, team and
column paths), sample rate, scale type (src/fcwt/fcwt.cpp:111-190), frequency range, number
of scales, Morlet sigma, thread count (which sets how the reference's shared
O1 buffer is reused between scales, fcwt.cpp:449-529), normalisation, the
scalar-coverage flag and real or complex input (fcwt.cpp:536-538), each
drawn per case from a fixed seed.  Sampled rows (the first, the last and
random ones) against the float64 oracle.  Seeded random batches of uneven
channels (fcwt_hip_batch_*) against one-signal plans, bit for bit, and the
input gradient (fcwt_hip_adjoint_device) of seeded random plans against the
float64 oracle adjoint.

Bar: relative L2 <= 1e-5 per row, and rows the oracle leaves all-zero are
zero.
"""
import numpy as np
import pytest

import fcwt_amd
from fcwt_amd import _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
CASES = 24


def _draw(case):
    rng = np.random.default_rng(9000 + case)
    n = int(np.exp(rng.uniform(np.log(64), np.log(600_000))))
    fs = int(rng.integers(100, 50_000))
    st = int(rng.integers(0, 3))
    lo, hi = sorted(np.exp(rng.uniform(np.log(1e-4), np.log(0.5), 2)))
    f0, f1 = lo * fs, max(hi * fs, lo * fs + 0.02)
    fn = int(rng.integers(1, 300))
    fb = float(rng.uniform(0.5, 12.0))
    threads = int(rng.choice([1, 2, 4, 8]))
    normalize = bool(rng.integers(0, 2))
    flags = nat.COVERAGE_SCALAR if rng.uniform() < 0.25 else 0
    complex_input = bool(rng.uniform() < 0.25)
    t = np.arange(n) / fs
    fr = rng.uniform(-0.45 if complex_input else 0.0, 0.45, 4) * fs
    if complex_input:
        x = np.exp(2j * np.pi * fr[:, None] * t[None, :]).sum(axis=0)
        x = (x + 0.3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)
    else:
        x = np.sin(2 * np.pi * fr[:, None] * t[None, :] + rng.uniform(0, 6.3, 4)[:, None]).sum(axis=0)
        x = (x + 0.3 * rng.standard_normal(n)).astype(np.float32)
    rows = sorted({0, fn - 1} | set(rng.integers(0, fn, min(fn, 6)).tolist()))
    return dict(n=n, fs=fs, st=st, f0=f0, f1=f1, fn=fn, fb=fb, threads=threads, normalize=normalize,
                flags=flags, complex_input=complex_input, x=x, rows=rows)


def _describe(c):
    return ", ".join(f"{k}={c[k]}" for k in ("n", "fs", "st", "f0", "f1", "fn", "fb", "threads", "normalize",
                                              "flags", "complex_input"))


@pytest.mark.parametrize("case", range(CASES))
def test_random_plan(case):
    import torch

    c = _draw(case)
    print(f"case {case}: {_describe(c)}", flush=True)
    sc = O.scales(c["st"], c["fs"], c["f0"], c["f1"], c["fn"])
    p = fcwt_amd.Plan(c["n"], sc, c["fb"], c["threads"], c["normalize"], c["flags"])
    dx = torch.from_numpy(c["x"]).to("cuda:0")
    out = torch.empty((c["fn"], c["n"]), dtype=torch.complex64, device="cuda:0")
    p.execute_device(dx.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream,
                     complex_input=c["complex_input"])
    torch.cuda.synchronize()
    p.check()
    got = out.cpu().numpy()
    assert np.all(np.isfinite(got))
    sel = c["rows"]
    avx = not (c["flags"] & nat.COVERAGE_SCALAR)
    ref = O.cwt_fp64(c["x"], sc, c["fb"], c["threads"], avx, c["normalize"], rows=sel)
    nz = np.linalg.norm(ref, axis=1) > 0
    assert np.all(got[sel][~nz] == 0), "rows the oracle leaves all-zero must be zero"
    err = O.rel_l2_rows(got[sel], ref)
    worst = int(np.argmax(np.where(nz, err, 0))) if nz.any() else 0
    assert err[nz].max(initial=0) <= TOL, f"row {sel[worst]}: rel-L2 {err[worst]:.3e} ({_describe(c)})"


def test_random_plans_cover_every_path():
    """The drawn cases reach the direct, team and column row paths (the
    column path's rows count as two-pass rows) and both input kinds (plan
    shapes only, no GPU work)."""
    kinds, paths = set(), {"direct": 0, "team": 0, "two_pass": 0}
    sizes = set()
    for case in range(CASES):
        c = _draw(case)
        sc = O.scales(c["st"], c["fs"], c["f0"], c["f1"], c["fn"])
        split = fcwt_amd.Plan(c["n"], sc, c["fb"], c["threads"], c["normalize"], c["flags"]).row_split()
        for k, v in split.items():
            paths[k] += v
        kinds.add(c["complex_input"])
        sizes.add(O.find2power(c["n"]))
    assert kinds == {False, True}
    assert min(paths.values()) > 0, paths
    assert max(sizes) >= 19 and min(sizes) <= 10, sorted(sizes)


BATCHES = 6


def _draw_batch(case):
    rng = np.random.default_rng(9500 + case)
    cnt = int(rng.integers(3, 11))
    fs = int(rng.integers(100, 50_000))
    fb = float(rng.uniform(0.5, 12.0))
    threads = int(rng.choice([1, 2, 4, 8]))
    normalize = bool(rng.integers(0, 2))
    flags = nat.COVERAGE_SCALAR if rng.uniform() < 0.25 else 0
    lengths, xs, scs = [], [], []
    for _ in range(cnt):
        n = int(np.exp(rng.uniform(np.log(64), np.log(300_000))))
        lo, hi = sorted(np.exp(rng.uniform(np.log(1e-4), np.log(0.5), 2)))
        sc = O.scales(int(rng.integers(0, 3)), fs, lo * fs, max(hi * fs, lo * fs + 0.02), int(rng.integers(1, 120)))
        t = np.arange(n) / fs
        fr = rng.uniform(0.0, 0.45, 3) * fs
        x = np.sin(2 * np.pi * fr[:, None] * t[None, :]).sum(axis=0) + 0.3 * rng.standard_normal(n)
        lengths.append(n)
        xs.append(x.astype(np.float32))
        scs.append(sc)
    return dict(lengths=lengths, xs=xs, scs=scs, fb=fb, threads=threads, normalize=normalize, flags=flags)


@pytest.mark.parametrize("case", range(BATCHES))
def test_random_batch(case):
    """A seeded batch of 3 .. 10 channels of random length (64 .. 300,000),
    each with its own random scales, one sigma, thread count, normalisation
    and coverage flag for the batch: every channel equals, bit for bit, the
    one-signal plan of the same arguments, and the first and last row of
    every channel meet the bar against the float64 oracle."""
    import torch

    c = _draw_batch(case)
    print(f"batch {case}: lengths {c['lengths']}, fns {[s.size for s in c['scs']]}, fb={c['fb']}, "
          f"threads={c['threads']}, normalize={c['normalize']}, flags={c['flags']}", flush=True)
    args = (c["fb"], c["threads"], c["normalize"], c["flags"])
    b = fcwt_amd.Batch(c["lengths"], c["scs"], *args)
    dx = [torch.from_numpy(x).cuda() for x in c["xs"]]
    outs = [torch.empty((s.size, x.size), dtype=torch.complex64, device="cuda") for s, x in zip(c["scs"], c["xs"])]
    b.execute_device([d.data_ptr() for d in dx], [o.data_ptr() for o in outs],
                     torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    b.check()
    avx = not (c["flags"] & nat.COVERAGE_SCALAR)
    for k, (x, sc, o) in enumerate(zip(c["xs"], c["scs"], outs)):
        got = o.cpu().numpy()
        p = fcwt_amd.Plan(x.size, sc, *args)
        single = torch.empty_like(o)
        p.execute_device(dx[k].data_ptr(), single.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        p.check()
        assert np.array_equal(got.view(np.uint64), single.cpu().numpy().view(np.uint64)), f"channel {k} (n={x.size})"
        rows = sorted({0, sc.size - 1})
        ref = O.cwt_fp64(x, sc, c["fb"], c["threads"], avx, c["normalize"], rows=rows)
        nz = np.linalg.norm(ref, axis=1) > 0
        assert np.all(got[rows][~nz] == 0)
        err = O.rel_l2_rows(got[rows], ref)
        assert err[nz].max(initial=0) <= TOL, f"channel {k} (n={x.size}): rel-L2 {err.max():.3e}"


ADJOINTS = 8


@pytest.mark.parametrize("case", range(ADJOINTS))
def test_random_adjoint(case):
    """The input gradient (fcwt_hip_adjoint_device) of a seeded random plan
    (length 64 .. 60,000, up to 120 scales of a random type and range,
    random sigma, thread count, normalisation, coverage flag and input kind)
    for a random output gradient, against the float64 oracle adjoint:
    relative L2 <= 1e-5 of the gradient vector."""
    import torch

    rng = np.random.default_rng(9800 + case)
    n = int(np.exp(rng.uniform(np.log(64), np.log(60_000))))
    fs = int(rng.integers(100, 50_000))
    lo, hi = sorted(np.exp(rng.uniform(np.log(1e-4), np.log(0.5), 2)))
    sc = O.scales(int(rng.integers(0, 3)), fs, lo * fs, max(hi * fs, lo * fs + 0.02), int(rng.integers(1, 120)))
    fb = float(rng.uniform(0.5, 12.0))
    threads = int(rng.choice([1, 2, 4, 8]))
    normalize = bool(rng.integers(0, 2))
    flags = nat.COVERAGE_SCALAR if rng.uniform() < 0.25 else 0
    complex_input = bool(rng.uniform() < 0.3)
    print(f"adjoint {case}: n={n}, fn={sc.size}, fb={fb}, threads={threads}, normalize={normalize}, "
          f"flags={flags}, complex_input={complex_input}", flush=True)
    G = (rng.standard_normal((sc.size, n)) + 1j * rng.standard_normal((sc.size, n))).astype(np.complex64)
    p = fcwt_amd.Plan(n, sc, fb, threads, normalize, flags)
    g = torch.from_numpy(G).cuda()
    xbar = torch.empty(n, dtype=torch.complex64 if complex_input else torch.float32, device="cuda")
    p.adjoint_device(g.data_ptr(), xbar.data_ptr(), torch.cuda.current_stream().cuda_stream, complex_input)
    torch.cuda.synchronize()
    got = xbar.cpu().numpy().astype(np.complex128)
    ref = O.adjoint_fp64(G, sc, fb, threads, not (flags & nat.COVERAGE_SCALAR), normalize, complex_input)
    ref = np.asarray(ref, np.complex128)
    if not np.linalg.norm(ref):
        assert not np.any(got)
        return
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= TOL, f"rel-L2 {err:.3e}"
