"""CPU FP64 oracle (test infrastructure only; never imported by the product package)."""
from .ravesim_oracle import *  # noqa: F401,F403
