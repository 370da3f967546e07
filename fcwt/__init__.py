"""The reference's Python package name (src/fcwt/__init__.py:1-2), so that
``import fcwt`` and ``from fcwt import *`` work unchanged: the same names,
served by fcwt_amd on the GPU."""
from fcwt_amd.api import FCWT, FCWT_LINFREQS, FCWT_LINSCALES, FCWT_LOGSCALES, Morlet, Scales
from fcwt_amd.boilerplate import cwt, plot

from . import boilerplate, fcwt  # noqa: F401  (the reference's submodules)

__all__ = ["Morlet", "Scales", "FCWT", "FCWT_LINSCALES", "FCWT_LOGSCALES", "FCWT_LINFREQS", "cwt", "plot"]
