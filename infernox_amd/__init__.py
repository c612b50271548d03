"""MI355X (gfx950) per-window search DP for infernox: p7 filter cascade
(F1 SSV, F1b MSV, F2 Viterbi, F3 Forward, F3b bias) and the covariance-model
scanners, behind the C ABI in include/ixg.h."""
from . import ixg  # noqa: F401
