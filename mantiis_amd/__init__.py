"""mantiis_amd -- MI355X (gfx950) D2Q9 lattice-Boltzmann step of mantiis.

The compute path is libmantiis_hip.so (hand-written HIP kernels behind the C ABI
of include/mantiis_hip.h).  This package holds its build script, a ctypes
binding used by the Python tests and bench.py, and the synthetic-geometry
generator.  The C++ drop-in classes live in include/*.hpp.
"""
from .hip import Lattice, HipLibraryMissing, load_library, MtsError  # noqa: F401
