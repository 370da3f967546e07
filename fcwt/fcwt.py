"""``fcwt.fcwt``: the SWIG module's classes (src/fcwt/fcwt.py), from fcwt_amd."""
from fcwt_amd.api import FCWT, FCWT_LINFREQS, FCWT_LINSCALES, FCWT_LOGSCALES, Morlet, Scales, Wavelet  # noqa: F401
