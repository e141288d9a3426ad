"""pyxtal_ff_amd -- the descriptor hot path of PyXtal_FF on MI355X (gfx950).

SO(4)/SNAP bispectrum and SO(3) power spectrum with atomic derivatives, a
device neighbour list, and a hand-written adjoint (energy/force/virial) path,
all in FP64 HIP kernels behind a C ABI (include/pxff.h).
"""
__version__ = "0.1.0"
