"""Drop-in ``SO4_Bispectrum`` with the SNAP conventions.

Same constructor, validation, ``calculate`` contract and output dict as the
reference class ``pyxtal_ff/descriptors/SNAP.py:10-334``: neighbour weights
from the ``weights`` dict keyed by element symbol, self weight 1.0,
psi = rfac0 * pi * r / rcut.  The numerics run on the GPU (libpxff.so); there
is no CPU path.
"""
import numpy as np

from ..structure import as_arrays
from ._so4_base import CUTOFF_NAMES, run_so4


class SO4_Bispectrum:
    """SNAP-convention SO(4) bispectrum (reference SNAP.py:10-63)."""

    def __init__(self, weights, lmax=3, rcut=3.5, derivative=True, stress=False, normalize_U=False,
                 cutoff_function='cosine', rfac0=0.99363):
        self.weights = weights
        self.lmax = lmax
        self.rcut = rcut
        self.derivative = derivative
        self.stress = stress
        self.normalize_U = normalize_U
        self.cutoff_function = cutoff_function
        self.rfac0 = rfac0
        self._type = "SO4"

    def __str__(self):
        s = "SO4 bispectrum descriptor with Cutoff: {:6.3f}".format(self.rcut)
        s += " lmax: {:d}\n".format(self.lmax)
        return s

    def __repr__(self):
        return str(self)

    def load_from_dict(self, dict0):
        self.lmax = dict0["lmax"]
        self.rcut = dict0["rcut"]
        self.normalize_U = dict0["normalize_U"]
        self.cutoff_function = dict0["cutoff_function"]
        self.derivative = dict0["derivative"]
        self.stress = dict0["stress"]

    def save_dict(self):
        return {"lmax": self.lmax, "rcut": self.rcut, "normalize_U": self.normalize_U,
                "cutoff_function": self.cutoff_function, "derivative": self.derivative,
                "stress": self.stress, "_type": "SO4"}

    # ---- validated properties (SNAP.py:65-187) ----
    @property
    def weights(self):
        return self._weights

    @weights.setter
    def weights(self, w):
        if isinstance(w, dict) is True:
            self._weights = w
        else:
            raise TypeError('weights must be a dictionary')

    @property
    def lmax(self):
        return self._twol // 2

    @lmax.setter
    def lmax(self, lmax):
        if isinstance(lmax, int) is True or isinstance(lmax, float) is True:
            if lmax < 0:
                raise ValueError('lmax must be greater than or equal to zero')
            elif lmax > 32:
                raise NotImplementedError('lmax > 32 is not supported')
            self._twol = round(2 * lmax)
        else:
            raise ValueError('lmax must be an integer')

    @property
    def rcut(self):
        return self._rcut

    @rcut.setter
    def rcut(self, rcut):
        if isinstance(rcut, float) is True or isinstance(rcut, int) is True:
            if rcut <= 0:
                raise ValueError('rcut must be greater than zero')
            self._rcut = rcut
        else:
            raise ValueError('rcut must be a float')

    @property
    def derivative(self):
        return self._derivative

    @derivative.setter
    def derivative(self, derivative):
        if isinstance(derivative, bool) is True:
            self._derivative = derivative
        else:
            raise ValueError('derivative must be a boolean value')

    @property
    def stress(self):
        return self._stress

    @stress.setter
    def stress(self, stress):
        if isinstance(stress, bool) is True:
            self._stress = stress
        else:
            raise ValueError('stress must be a boolean value')

    @property
    def normalize_U(self):
        return self._norm

    @normalize_U.setter
    def normalize_U(self, normalize_U):
        if isinstance(normalize_U, bool) is True:
            self._norm = normalize_U
        else:
            raise ValueError('normalize_U must be a boolean value')

    @property
    def cutoff_function(self):
        return self._cutoff_function

    @cutoff_function.setter
    def cutoff_function(self, cutoff_function):
        if isinstance(cutoff_function, str) is True:
            if cutoff_function in CUTOFF_NAMES:
                self._cutoff_id = CUTOFF_NAMES.index(cutoff_function) + 1
                self._cutoff_function = cutoff_function
            else:
                raise NotImplementedError('The requested cutoff function has not been implemented')
        else:
            raise ValueError('You must specify the cutoff function as a string')

    @property
    def rfac0(self):
        return self._rfac0

    @rfac0.setter
    def rfac0(self, rfac0):
        if isinstance(rfac0, float) is True or isinstance(rfac0, int) is True:
            if rfac0 <= 0:
                raise ValueError('rfac0 must be greater than zero')
            elif rfac0 > 1:
                raise ValueError('rfac0 must be less than or equal to one')
            self._rfac0 = rfac0
        else:
            raise ValueError('rfac0 must be a float')

    def clear_memory(self):
        attrs = list(vars(self).keys())
        for attr in attrs:
            if attr not in {'_twol', '_rcut', '_derivative', '_stress', '_norm', '_cutoff_function', '_cutoff_id',
                            '_weights', '_rfac0', '_type'}:
                delattr(self, attr)

    def _weights_of(self, atoms):
        _, _, numbers, symbols, _ = as_arrays(atoms)
        aw = np.array([self.weights[s] for s in symbols], dtype=np.float64)
        return aw, np.ones(len(symbols))

    def calculate(self, atoms, atom_ids=None, as_numpy=True):
        """dict with 'x' [N, nc], 'dxdr' [npairs, nc, 3], 'rdxdr' [npairs, nc,
        3, 3] (= -bstress/V per seq row), 'elements', 'seq' [npairs, 2]
        (SNAP.py:200-238)."""
        aw, sw = self._weights_of(atoms)
        x = run_so4(atoms, self._twol, self.rcut, self.rfac0, self._cutoff_id, self._norm, self.derivative,
                    self.stress, aw, sw, atom_ids=atom_ids, as_numpy=as_numpy)
        self.clear_memory()
        return x
