"""fcwt_amd -- fCWT's Morlet ``FCWT::cwt`` on AMD MI355X (gfx950).

Drop-in for the reference's Python package: ``import fcwt_amd as fcwt`` gives
``Morlet``, ``Scales``, ``FCWT``, ``FCWT_LINSCALES/LOGSCALES/LINFREQS``,
``cwt`` and ``plot`` with the reference's signatures (src/fcwt/__init__.py:1-2).
"""
from .api import FCWT, FCWT_LINFREQS, FCWT_LINSCALES, FCWT_LOGSCALES, Batch, Morlet, Plan, Scales, Wavelet
from .boilerplate import cwt, plot

__all__ = ["Morlet", "Scales", "FCWT", "Plan", "Batch", "Wavelet", "FCWT_LINSCALES", "FCWT_LOGSCALES", "FCWT_LINFREQS",
           "cwt", "plot"]
