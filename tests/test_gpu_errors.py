"""Failure handling on the GPU, with every failure forced from the host so
that every kernel still exits normally: a device-memory budget that plan
creation exceeds (the allocation-failure path), and a spin limit of one poll
on the wide-band kernel's own bounded waits (the timed-out-wait path).  After
each failure the library must report it once, hold no half-built plan, and
compute correctly again."""
import numpy as np
import pytest

import fcwt_amd
from fcwt_amd import _native as nat
from oracle import oracle as O
from tests.gpu_util import run_device, tones_noise

pytestmark = pytest.mark.gpu


def test_allocation_failure_is_reported_and_recovers(monkeypatch):
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 3000)
    monkeypatch.setenv("FCWT_HIP_DEVICE_BUDGET_MB", "1")
    with pytest.raises(nat.FcwtError) as e:
        fcwt_amd.Plan(100_000, sc, 2.0)
    assert e.value.code == nat.ENOMEM and "budget" in str(e.value)
    # a batch (one plan per padded size) fails the same way
    with pytest.raises(nat.FcwtError):
        fcwt_amd.Batch([50_000, 60_000], sc[:40], 2.0)
    monkeypatch.delenv("FCWT_HIP_DEVICE_BUDGET_MB")
    x = tones_noise(100_000, seed=3)
    rows = [0, 1000, 2999]
    got = run_device(x, sc, 2.0)[rows]
    assert O.rel_l2_rows(got, O.cwt_fp64(x, sc, 2.0, rows=rows)).max() <= 1e-5


def test_timed_out_wait_is_reported_once_and_recovers(monkeypatch):
    """A spin limit of zero polls: the first team wait of the wide-band
    kernel fails, records the error, and every work-group exits;
    plan_check reports it, and a later check is clean."""
    import torch

    x = tones_noise(100_000, seed=4)
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 300)
    monkeypatch.setenv("FCWT_HIP_TEAM_SPIN", "0")
    p = fcwt_amd.Plan(x.size, sc, 2.0)
    assert p.row_split()["team"] > 0
    dx = torch.from_numpy(x).cuda()
    out = torch.empty((sc.size, x.size), dtype=torch.complex64, device="cuda")
    p.execute_device(dx.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    with pytest.raises(nat.FcwtError) as e:
        p.check()
    assert "timed out" in str(e.value)
    p.check()  # reported once; the error word is clear again
    monkeypatch.delenv("FCWT_HIP_TEAM_SPIN")
    got = run_device(x, sc, 2.0)
    rows = [0, 150, 299]
    assert O.rel_l2_rows(got[rows], O.cwt_fp64(x, sc, 2.0, rows=rows)).max() <= 1e-5


def test_library_before_torch_shares_one_runtime():
    """A fresh process that calls the library before it touches torch: torch
    bundles a HIP runtime under the same soname as /opt/rocm's, and the one
    loaded first serves both, so loading the library must not leave torch
    without a GPU.  The child runs fcwt.cwt on host arrays, then a torch
    kernel, and checks the transform against the oracle."""
    import os
    import subprocess
    import sys

    code = (
        "import numpy as np, fcwt_amd\n"
        "from oracle import oracle as O\n"
        "from tests.gpu_util import tones_noise\n"
        "x = tones_noise(4000, fs=100, fmax=45, seed=3)\n"
        "freqs, y = fcwt_amd.cwt(x, 100, 1, 40, 12)\n"
        "import torch\n"
        "t = torch.arange(8, dtype=torch.float32, device='cuda:0')\n"
        "assert float(t.sum().item()) == 28.0\n"
        "sc = O.scales(O.LINFREQS, 100, 1, 40, 12)\n"
        "err = O.rel_l2_rows(y, O.cwt_fp64(x, sc, 2.0)).max()\n"
        "assert err <= 1e-5, err\n"
        "print('ok', err)\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]


def test_failed_launch_is_not_sticky_and_surfaces_through_autograd(monkeypatch):
    """One forced failure on a plan that stays in use: the failed launch's
    wide-band rows are NaN (never stale values that look valid), check()
    raises once, and the next forward of the same module and plan -- one
    signal through the plan, two through the batch plan -- is correct."""
    import torch

    from fcwt_amd.autograd import CWT

    n = 60_000
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 200)
    op = CWT(n, sc, 2.0)
    split = op.plan.row_split()
    assert split["team"] > 0 and split["direct"] > 0
    x = torch.from_numpy(np.stack([tones_noise(n, seed=s) for s in (5, 6)])).cuda()
    monkeypatch.setenv("FCWT_HIP_TEAM_SPIN", "0")
    for xin in (x[0], x):
        W = op(xin)
        torch.cuda.synchronize()
        with pytest.raises(nat.FcwtError) as e:
            op.check()
        assert "timed out" in str(e.value)
        finite = torch.isfinite(W.reshape(-1, sc.size, n)).all(dim=-1).cpu().numpy()
        assert not finite.all(axis=1).any()  # every signal has NaN rows (its team rows)
        assert finite.any()                  # the direct rows were still written
        op.check()                           # reported once
    monkeypatch.delenv("FCWT_HIP_TEAM_SPIN")
    rows = [0, 60, 120, 199]
    for xin in (x[0], x):
        W = op(xin).reshape(-1, sc.size, n)
        op.check()
        for b in range(W.shape[0]):
            got = W[b].cpu().numpy()
            assert np.isfinite(got).all()
            ref = O.cwt_fp64(x[b].cpu().numpy(), sc, 2.0, rows=rows)
            assert O.rel_l2_rows(got[rows], ref).max() <= 1e-5
    # opt-in checking after every forward raises at the call itself
    op.check_errors = True
    monkeypatch.setenv("FCWT_HIP_TEAM_SPIN", "0")
    with pytest.raises(nat.FcwtError):
        op(x[0])
    monkeypatch.delenv("FCWT_HIP_TEAM_SPIN")
    op(x[0])
