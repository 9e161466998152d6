"""kore_amd: MI355X (gfx950) solver backend for Kore's solve path.

Modules
-------
params    Kore's solver/physics knobs and derived sizes (bin/parameters.py, bin/utils.py)
assembly  A, B assembly for the hydrodynamic case (input side of the path)
backend   ctypes binding of the HIP library libkore_amd.so (include/kore_amd.h)
eigs      shift-and-invert block Krylov-Schur on the device
solve     drop-in for bin/solve.py
serving   request loop with continuous batching over resident factors
"""
__version__ = "0.1.0"
