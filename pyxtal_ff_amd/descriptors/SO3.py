"""Drop-in ``SO3`` power spectrum descriptor.

Same constructor, validation, ``calculate`` contract and output dict as the
reference class ``pyxtal_ff/descriptors/SO3.py:8-315``; the numerics run on
the GPU (libpxff.so), there is no CPU path.
"""
import numpy as np
import torch

from .. import so3 as dev_so3
from .. import so4 as dev_so4
from ..neighbors import DeviceNeighborList
from ..structure import as_arrays
from ._so4_base import prefix_count

CUTOFFS = ("cosine", "tanh", "poly1", "poly2", "poly3", "poly4", "exp", "unity")


class SO3:
    """SO(3) power spectrum of the Gaussian neighbour density (SO3.py:8-66)."""

    def __init__(self, nmax=3, lmax=3, rcut=3.5, alpha=2.0, derivative=True, stress=False, cutoff_function='cosine',
                 weight_on=False, primitive=False):
        self.nmax = nmax
        self.lmax = lmax
        self.rcut = rcut
        self.alpha = alpha
        self.derivative = derivative
        self.stress = stress
        self._type = "SO3"
        self.cutoff_function = cutoff_function
        self.weight_on = weight_on
        self.primitive = primitive

    def __str__(self):
        s = "SO3 descriptor with Cutoff: {:6.3f}".format(self.rcut)
        s += " lmax: {:d}, nmax: {:d}, alpha: {:.3f}\n".format(self.lmax, self.nmax, self.alpha)
        return s

    def __repr__(self):
        return str(self)

    def load_from_dict(self, dict0):
        self.nmax = dict0["nmax"]
        self.lmax = dict0["lmax"]
        self.rcut = dict0["rcut"]
        self.alpha = dict0["alpha"]
        self.derivative = dict0["derivative"]
        self.stress = dict0["stress"]

    def save_dict(self):
        return {"nmax": self.nmax, "lmax": self.lmax, "rcut": self.rcut, "alpha": self.alpha,
                "derivative": self.derivative, "stress": self.stress, "_type": "SO3"}

    # ---- validated properties (SO3.py:68-175) ----
    @property
    def nmax(self):
        return self._nmax

    @nmax.setter
    def nmax(self, nmax):
        if isinstance(nmax, int) is True:
            if nmax < 1:
                raise ValueError('nmax must be greater than or equal to 1')
            if nmax > 11:
                raise ValueError('nmax > 11 yields complex eigenvalues which will mess up the calculation')
            self._nmax = nmax
        else:
            raise ValueError('nmax must be an integer')

    @property
    def lmax(self):
        return self._lmax

    @lmax.setter
    def lmax(self, lmax):
        if isinstance(lmax, int) is True:
            if lmax < 0:
                raise ValueError('lmax must be greater than or equal to zero')
            elif lmax > 32:
                raise NotImplementedError('lmax > 32 is not supported')
            self._lmax = lmax
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
    def alpha(self):
        return self._alpha

    @alpha.setter
    def alpha(self, alpha):
        if isinstance(alpha, float) is True or isinstance(alpha, int) is True:
            if alpha <= 0:
                raise ValueError('alpha must be greater than zero')
            self._alpha = alpha
        else:
            raise ValueError('alpha must be a float')

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
    def cutoff_function(self):
        return self._cutoff_function

    @cutoff_function.setter
    def cutoff_function(self, cutoff_function):
        if isinstance(cutoff_function, str) is True:
            if cutoff_function in CUTOFFS:
                self._cutoff_function = cutoff_function
            else:
                raise NotImplementedError('The requested cutoff function has not been implemented')
        else:
            raise ValueError('You must specify the cutoff function as a string')

    def clear_memory(self):
        attrs = list(vars(self).keys())
        for attr in attrs:
            if attr not in {'_nmax', '_lmax', '_rcut', '_alpha', '_derivative', '_stress', '_cutoff_function',
                            'weight_on', 'primitive', '_type'}:
                delattr(self, attr)

    def calculate(self, atoms, atom_ids=None, as_numpy=True):
        """dict with 'x' [N, ncoef], 'dxdr' [nseq, ncoef, 3], 'rdxdr' [nseq,
        ncoef, 3, 3] (-pstress/V), 'elements', 'seq' (SO3.py:187-291)."""
        if self.derivative and self._cutoff_function == "exp":
            # SO3.py:442-457: Exponent(derivative=True) returns None, so the
            # reference cannot differentiate this cutoff
            raise NotImplementedError("the 'exp' cutoff has no derivative in the reference SO3")
        pos, cell, numbers, symbols, pbc = as_arrays(atoms)
        n = len(pos)
        k = prefix_count(atom_ids, n)
        nl = DeviceNeighborList([pos], [cell], [pbc], self.rcut).resolve()  # count settled before any kernel
        cfg = dev_so3.So3Config(self.nmax, self.lmax, self.rcut, self.alpha, self._cutoff_function, self.weight_on)
        d = nl.pos.device
        z = torch.as_tensor(numbers, dtype=torch.int64, device=d)
        if self.derivative:
            x, ct, rad = dev_so3.forward(nl, cfg, z, radial=True)
        else:
            x, ct = dev_so3.forward(nl, cfg, z)
        x = x.clone()
        if k < n:
            x[k:] = 0.0  # the reference leaves rows of atoms outside atom_ids at zero
        out = {"x": x}
        if self.derivative:
            dp = dev_so3.dpdr(nl, cfg, z, ct, radial=rad)
            dx, rd = dev_so4.rows(nl, cfg.ncoef, dp, stress=self.stress)
            nr = int(nl.row_start[k].item())
            out["dxdr"] = dx[:nr]
            out["elements"] = list(symbols)
            out["seq"] = nl.seq_array()[:nr]
            out["rdxdr"] = rd[:nr] if self.stress else None
        else:
            out.update({"dxdr": None, "rdxdr": None, "elements": list(symbols)})
        if as_numpy:
            for key in ("x", "dxdr", "rdxdr", "seq"):
                if isinstance(out.get(key), torch.Tensor):
                    out[key] = out[key].cpu().numpy()
        self.clear_memory()
        return out
