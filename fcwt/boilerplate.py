"""``fcwt.boilerplate`` (src/fcwt/boilerplate.py): ``cwt`` and ``plot``, from fcwt_amd."""
from fcwt_amd.boilerplate import cwt, plot  # noqa: F401
