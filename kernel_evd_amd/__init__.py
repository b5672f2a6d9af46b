"""kernel_evd_amd -- MI355X (gfx950) accelerated randomly pivoted Cholesky,
bistochastic kernel EVD and bandwidth sweep, behind the C ABI of
include/kevd.h (libkevd.so in this directory)."""
from .api import Context, ArpcOutput, EvdOutput, bandwidth_post, delay_embed, ks_generate  # noqa: F401
from ._lib import KevdError, LIB_PATH  # noqa: F401
