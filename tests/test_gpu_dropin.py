"""The Python drop-in on the GPU: ``fcwt.cwt`` keeps its device plan between
calls of the same shape, and ``fcwt.plot`` (reference boilerplate.py:81-88,
with the arity of its ``_plot`` call fixed) runs the transform and draws it."""
import numpy as np
import pytest

import fcwt
from fcwt_amd import boilerplate as bp
from oracle import oracle as O
from tests.gpu_util import tones_noise

pytestmark = pytest.mark.gpu


def test_cwt_reuses_its_plan_between_calls():
    bp.clear_plan_cache()
    x = tones_noise(20_000, seed=21)
    f1, out1 = fcwt.cwt(x, 1000, 1, 450, 120, scaling="log")
    plan = bp._FCWT_CACHE[(1, False, True)]._plan
    f2, out2 = fcwt.cwt(x, 1000, 1, 450, 120, scaling="log")
    assert bp._FCWT_CACHE[(1, False, True)]._plan is plan
    assert np.array_equal(out1, out2) and np.array_equal(f1, f2)
    # another shape builds another plan, and the result stays right
    y = tones_noise(12_345, seed=22)
    _, out3 = fcwt.cwt(y, 1000, 1, 450, 64)
    assert bp._FCWT_CACHE[(1, False, True)]._plan is not plan
    sc = O.scales(O.LINFREQS, 1000, 1, 450, 64)
    rows = [0, 31, 63]
    assert O.rel_l2_rows(out3[rows], O.cwt_fp64(y, sc, 2.0, rows=rows)).max() <= 1e-5
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 120)
    assert O.rel_l2_rows(out1[rows], O.cwt_fp64(x, sc, 2.0, rows=rows)).max() <= 1e-5
    bp.clear_plan_cache()
    assert not bp._FCWT_CACHE


def test_plot_runs_the_transform(monkeypatch):
    mpl = pytest.importorskip("matplotlib")
    mpl.use("Agg")
    import matplotlib.pyplot as plt

    shown = []
    monkeypatch.setattr(plt, "show", lambda: shown.append(plt.gcf()))
    fs = 100
    x = np.sin(2 * np.pi * 7 * np.arange(6000) / fs).astype(np.float32)
    fcwt.plot(x, fs, f0=1, f1=40, fn=50)
    assert len(shown) == 1
    img = shown[0].axes[1].images[0].get_array()
    assert img.shape == (50, 6000) and np.isfinite(img).all()
    # the peak row is the one nearest 7 Hz (LINFREQS 1..40, fcwt.cpp:177-190)
    freqs = O.frequencies(O.scales(O.LINFREQS, fs, 1, 40, 50), fs)
    peak = int(np.argmax(np.asarray(img)[:, 1000:5000].mean(axis=1)))
    assert abs(freqs[peak] - 7) < 1.0, freqs[peak]
    plt.close("all")
