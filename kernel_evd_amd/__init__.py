"""kernel_evd_amd -- MI355X (gfx950) accelerated randomly pivoted Cholesky,
bistochastic kernel EVD and bandwidth sweep, behind the C ABI of
include/kevd.h (libkevd.so in this directory)."""
from .api import (Context, ArpcOutput, EvdOutput, bandwidth_post, delay_embed, ensemble_arpc,  # noqa: F401
                  ensemble_evd, ks_generate)
from ._lib import KevdError, LIB_PATH  # noqa: F401
