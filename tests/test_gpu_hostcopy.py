"""The host-pointer entry point's pipelined output copy (fcwt_hip_cwt into a
caller's pageable array, through two pinned 64 MiB staging buffers) against
the device-pointer output of the same plan, bit for bit, and against one
plain copy (FCWT_HIP_NO_PIPELINE=1).  The reference writes its output into
the caller's host array directly (src/fcwt/fcwt.cpp:514-518, 532-548)."""
import numpy as np
import pytest

import fcwt_amd
from oracle import oracle as O
from tests.gpu_util import tones_noise

pytestmark = pytest.mark.gpu

STAGE_BYTES = 64 << 20


def _device_output(plan, x):
    import torch

    dx = torch.from_numpy(x).cuda()
    out = torch.empty((plan.fn, plan.n), dtype=torch.complex64, device="cuda")
    plan.execute_device(dx.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    plan.check()
    host = out.cpu().numpy()
    del out
    torch.cuda.empty_cache()
    return host


@pytest.mark.parametrize("n,fn", [(2_000_003, 40), (100_000, 3000)])
def test_pipelined_host_copy_matches_device_output(monkeypatch, n, fn):
    """n = 2,000,003 x 40 rows is 640 MB: ten chunks, the last one partial
    and rows split across chunk edges.  100k x 3000 is the 2.4 GB headline
    output: 36 chunks."""
    fs = 1000
    x = tones_noise(n, fs=fs, seed=17)
    sc = O.scales(O.LOGSCALES, fs, 1, 450, fn)
    plan = fcwt_amd.Plan(n, sc, 2.0)
    nbytes = 8 * n * fn
    assert nbytes > 3 * STAGE_BYTES and nbytes % STAGE_BYTES != 0
    ref_dev = _device_output(plan, x)
    monkeypatch.delenv("FCWT_HIP_NO_PIPELINE", raising=False)
    got = np.full((fn, n), np.nan, np.complex64)
    plan.execute_host(x, got)
    assert np.array_equal(got.view(np.uint64), ref_dev.view(np.uint64))
    # a second call into the same array reuses the staging buffers
    got2 = np.full((fn, n), np.nan, np.complex64)
    plan.execute_host(x, got2)
    assert np.array_equal(got2.view(np.uint64), ref_dev.view(np.uint64))
    del got2
    monkeypatch.setenv("FCWT_HIP_NO_PIPELINE", "1")
    plain = np.full((fn, n), np.nan, np.complex64)
    plan.execute_host(x, plain)
    assert np.array_equal(plain.view(np.uint64), ref_dev.view(np.uint64))
    del plain
    # rows on either side of the chunk edges, against the oracle
    row_bytes = 8 * n
    edge_rows = sorted({min(fn - 1, (k * STAGE_BYTES) // row_bytes) for k in range(1, nbytes // STAGE_BYTES + 1)})
    rows = sorted(set(edge_rows[:4] + [0, fn - 1]))
    ref = O.cwt_fp64(x, sc, 2.0, rows=rows)
    assert O.rel_l2_rows(got[rows], ref).max() <= 1e-5
    plan.close()
