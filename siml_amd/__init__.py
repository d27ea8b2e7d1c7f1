"""siml_amd -- MI355X-native LINGO chemical similarity (the hot path of
pySIML 1.5) behind the reference's own interface.

Modules mirror the reference package:
  siml_amd.compiler   cSMILEStoMatrices / SMILEStoMatrices (host packer)
  siml_amd.GPULingo   GPULingo comparator (HIP kernels on gfx950)
  siml_amd.CPULingo   CPULingo interface, served by the same kernels
  siml_amd.utilities  rowHistogram101
The compute path is libsiml_amd.so (C ABI in include/siml_amd.h); importing
a comparator without the built library raises ImportError.
"""
__version__ = "1.5+mi355x.1"

from . import _lib  # noqa: F401


def build(verbose=False):
    """Compile libsiml_amd.so in-tree for gfx950."""
    from . import _build
    return _build.build(verbose=verbose)
