"""gvmps -- MI355X-native MPS simulator backend (observe / sample / gradient).

Host mirror of the reference's custom MPS NVQIR plugin call surface over the
native library libgvmps.so (include/gvmps.h).
"""
from .engine import Batch, Problem, encode_ops  # noqa: F401
