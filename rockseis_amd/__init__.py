"""rockseis_amd: MI355X-native staggered-grid FD time step and FWI gradient
behind RockSeis's Waves/Der/Pml/Fwi/Revolve interfaces.

The compute path is librsgpu.so (HIP kernels for gfx950 plus a C-ABI,
declared in include/rsgpu.h); the Python modules here call it through
ctypes.  Importing the package does not load the library; the first call
that needs it does, and fails loudly if it is missing.
"""
__version__ = "0.1.0"
