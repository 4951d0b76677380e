"""b2g: MI355X-native two-site DMRG effective-Hamiltonian hot path.

Python face of libb2g.so (C ABI in include/b2g.h): the block-sparse problem
container, a GPU EffectiveHamiltonian with block2's method names, and the
synthetic middle-site generator used by bench.py.
"""
from ._lib import LIB_PATH, B2GError, lib  # noqa: F401
from .problem import InfoSet, OpSet, Problem  # noqa: F401
