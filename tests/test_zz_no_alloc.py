"""fcwt_hip_cwt_device_noalloc through the C ABI: the same output, bit for
bit, as fcwt_hip_cwt_device on a plan of the same signal length and scales,
on two calls in a row (the second smaller, so the static array's plan is
rebuilt), without allocating device memory; sizes beyond its limit are
rejected with a clear error."""
import ctypes

import numpy as np
import pytest
import torch

import fcwt_amd
from fcwt_amd import _native as nat
from oracle import oracle as O
from tests.gpu_util import tones_noise

pytestmark = pytest.mark.gpu


def _noalloc(x_dev, n, sc, out_dev, fb=2.0, threads=1, normalize=1, flags=0):
    L = nat.lib()
    st = torch.cuda.current_stream().cuda_stream
    return L.fcwt_hip_cwt_device_noalloc(ctypes.c_void_p(x_dev.data_ptr()), n, nat.fptr(sc), sc.size, fb, threads,
                                         normalize, flags, ctypes.c_void_p(out_dev.data_ptr()), ctypes.c_void_p(st))


def _with_plan(x, sc, fb=2.0):
    plan = fcwt_amd.Plan(x.size, sc, fb)
    dx = torch.from_numpy(x).cuda()
    out = torch.empty((sc.size, x.size), dtype=torch.complex64, device="cuda")
    plan.execute_device(dx.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    plan.check()
    plan.close()
    return out.cpu().numpy()


@pytest.mark.parametrize("cases", [[(100_000, 3000), (30_000, 500)], [(131_072, 1000), (5_000, 64)]])
def test_noalloc_matches_plan_entry_point(cases):
    for n, fn in cases:
        x = tones_noise(n, seed=n)
        sc = O.scales(O.LOGSCALES, 1000, 1, 450, fn)
        ref = _with_plan(x, sc)
        dx = torch.from_numpy(x).cuda()
        out = torch.full((fn, n), float("nan"), dtype=torch.complex64, device="cuda")
        assert _noalloc(dx, n, sc, out) == nat.OK, nat.lib().fcwt_hip_last_error()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (n, fn)
        rows = [0, fn // 2, fn - 1]
        assert O.rel_l2_rows(got[rows], O.cwt_fp64(x, sc, 2.0, rows=rows)).max() <= 1e-5


def test_noalloc_allocates_nothing_and_rejects_large_sizes():
    n, fn = 60_000, 400
    x = tones_noise(n, seed=8)
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, fn)
    dx = torch.from_numpy(x).cuda()
    out = torch.empty((fn, n), dtype=torch.complex64, device="cuda")
    assert _noalloc(dx, n, sc, out) == nat.OK
    torch.cuda.synchronize()
    first = out.cpu().numpy()
    free0 = torch.cuda.mem_get_info()[0]
    for seed in (9, 10):  # same plan (reused), then new scales (rebuilt in the array)
        sc2 = sc if seed == 9 else O.scales(O.LOGSCALES, 1000, 2, 400, fn)
        assert _noalloc(dx, n, sc2, out) == nat.OK
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0
    assert _noalloc(dx, n, sc, out) == nat.OK
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), first)
    big = torch.zeros(200_000, device="cuda")
    big_out = torch.empty(1, dtype=torch.complex64, device="cuda")
    rc = _noalloc(big, 200_000, sc, big_out)
    assert rc == nat.EUNSUPPORTED
    assert b"131072" in nat.lib().fcwt_hip_last_error()
