"""sehf -- MI355X static-exchange Hartree-Fock photoionization solver (host side).

The compute path is the HIP library spion-emol_amd/_build/libsehf.so behind
the C ABI in include/sehf.h; this package is the Python side of that ABI
(ctypes), the synthetic input generator and the observables layer.
"""
from ._lib import LIB_PATH, SehfError, load  # noqa: F401
