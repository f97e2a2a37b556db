"""MI355X-native accelerator for smuthi's lookup-table particle-coupling solve
and plane-wave-expansion near field.

Python host code mirrors smuthi's interfaces (Simulation, LinearSystem,
CouplingMatrixRadialLookupCUDA, CouplingMatrixVolumeLookupCUDA,
PlaneWaveExpansion) and calls hand-written gfx950 HIP kernels in libsmh.so
through a C ABI (include/smh.h) with ctypes.
"""
__version__ = '0.1.0'
