"""Host-side Python binding of libdesalm (MI355X pseudo-time inner iteration
of DES-ALM).  The computation is in the HIP library; see include/desalm.h."""
from .params import Params, make_params, PREC_FP64, PREC_MIXED  # noqa: F401
from . import cases  # noqa: F401
