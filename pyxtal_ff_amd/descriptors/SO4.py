"""Drop-in ``SO4_Bispectrum`` with the SO4.py conventions.

Reference ``pyxtal_ff/descriptors/SO4.py:10-304``: neighbour weight = atomic
number Z_j, self weight Z_i, psi = pi * r / rcut, integer lmax only.
Identical interface and output dict; the numerics run on the GPU.
"""
import numpy as np

from ..structure import as_arrays
from . import SNAP as _snap
from ._so4_base import run_so4


class SO4_Bispectrum(_snap.SO4_Bispectrum):
    """Z-weighted SO(4) bispectrum (reference SO4.py:10-158)."""

    def __init__(self, lmax=3, rcut=3.5, derivative=True, stress=False, normalize_U=False,
                 cutoff_function='cosine'):
        self._weights = {}
        self._rfac0 = 1.0
        self.lmax = lmax
        self.rcut = rcut
        self.derivative = derivative
        self.stress = stress
        self.normalize_U = normalize_U
        self.cutoff_function = cutoff_function
        self._type = "SO4"

    @property
    def lmax(self):
        return self._twol // 2

    @lmax.setter
    def lmax(self, lmax):
        if isinstance(lmax, int) is True:
            if lmax < 0:
                raise ValueError('lmax must be greater than or equal to zero')
            elif lmax > 32:
                raise NotImplementedError('lmax > 32 is not supported')
            self._twol = 2 * lmax
        else:
            raise ValueError('lmax must be an integer')

    def _weights_of(self, atoms):
        _, _, numbers, _, _ = as_arrays(atoms)
        z = numbers.astype(np.float64)
        return z, z

    def calculate(self, atoms, atom_ids=None, as_numpy=True):
        aw, sw = self._weights_of(atoms)
        x = run_so4(atoms, self._twol, self.rcut, 1.0, self._cutoff_id, self._norm, self.derivative, self.stress,
                    aw, sw, atom_ids=atom_ids, as_numpy=as_numpy)
        self.clear_memory()
        return x
