"""The drop-in boundary without a GPU: the reference's own C++ example
compiles and links against include/fcwt.h + libfcwt_hip.so unchanged, the
`fcwt` package name serves `from fcwt import *`, and the plot helper draws."""
from __future__ import annotations

import pathlib
import re
import subprocess

import numpy as np
import pytest

from fcwt_amd import _native as nat

ROOT = pathlib.Path(__file__).resolve().parents[1]
REF_README = pathlib.Path("/root/reference/README.md")


def _readme_cpp_example():
    """The C++ block under the 'C++ Example' heading of the reference README
    (README.md:142-213), read where the reference lies."""
    text = REF_README.read_text()
    start = text.index("C++ Example")
    m = re.search(r"```cpp\n(.*?)```", text[start:], re.S)
    assert m, "no cpp block after 'C++ Example'"
    return m.group(1)


@pytest.mark.skipif(not REF_README.exists(), reason="reference sources not present")
def test_reference_readme_cpp_example_compiles(tmp_path):
    """README.md:142-213 uses complex<float> unqualified; it relies on the
    reference header's `using namespace std;` (fcwt.h:63)."""
    if not nat.LIB_PATH.exists():
        pytest.skip("libfcwt_hip.so not built")
    src = tmp_path / "readme_example.cpp"
    src.write_text(_readme_cpp_example())
    exe = tmp_path / "readme_example"
    r = subprocess.run(["g++", "-std=c++17", f"-I{ROOT / 'include'}", str(src), str(nat.LIB_PATH),
                        f"-Wl,-rpath,{nat.LIB_PATH.parent}", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert exe.exists()


def test_header_opens_std_like_the_reference(tmp_path):
    """A caller that writes vector and complex<float> unqualified compiles."""
    src = tmp_path / "unqualified.cpp"
    src.write_text('#include <vector>\n#include "fcwt.h"\n'
                   "int main() { vector<complex<float>> v(4); Morlet m(2.0f); (void)m; return (int)v.size() - 4; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{ROOT / 'include'}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_fcwt_package_star_import():
    """`from fcwt import *` (tests/test_fcwt.py:1) gives the names of
    src/fcwt/__init__.py:1-2, and they are fcwt_amd's."""
    import fcwt_amd

    ns = {}
    exec("from fcwt import *", ns)
    for name in ("Morlet", "Scales", "FCWT", "FCWT_LINSCALES", "FCWT_LOGSCALES", "FCWT_LINFREQS", "cwt", "plot"):
        assert ns[name] is getattr(fcwt_amd, name), name
    import fcwt

    assert fcwt.fcwt.Morlet is fcwt_amd.Morlet and fcwt.boilerplate.cwt is fcwt_amd.cwt
    assert (fcwt.FCWT_LINSCALES, fcwt.FCWT_LOGSCALES, fcwt.FCWT_LINFREQS) == (0, 1, 2)


def test_plot_helper_draws(monkeypatch):
    """boilerplate._plot (reference boilerplate.py:36-78) draws the input and
    |CWT| panels with the frequency and time ticks; show() is not called
    interactively under the Agg backend."""
    mpl = pytest.importorskip("matplotlib")
    mpl.use("Agg")
    import matplotlib.pyplot as plt

    from fcwt_amd import boilerplate as bp

    shown = []
    monkeypatch.setattr(plt, "show", lambda: shown.append(plt.gcf()))
    fs, n, fn = 100, 3000, 20
    x = np.sin(2 * np.pi * 5 * np.arange(n) / fs).astype(np.float32)
    freqs = np.linspace(50, 1, fn).astype(np.float32)
    out = (np.random.default_rng(0).standard_normal((fn, n)) + 0j).astype(np.complex64)
    bp._plot(x, freqs, out, fs, 1, 50, fn)
    assert len(shown) == 1
    ax1, ax2 = shown[0].axes
    assert ax1.get_title() == "Input signal" and ax2.get_title() == "CWT"
    assert np.allclose(ax2.images[0].get_array(), np.abs(out))
    labels = [t.get_text() for t in ax2.get_yticklabels()]
    assert labels[0] == "50.0" and labels[-1] == "1.0"
    plt.close("all")
