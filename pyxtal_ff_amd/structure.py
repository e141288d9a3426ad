"""Minimal structure type with the ASE ``Atoms`` surface the descriptors use.

The drop-in descriptor classes accept any object exposing ``positions``,
``cell``, ``numbers``, ``pbc`` and ``get_volume()`` (an ASE ``Atoms`` object
qualifies); ASE itself is not required.  ``Structure`` is that surface.
"""
import numpy as np

SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl",
           "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se",
           "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb",
           "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er",
           "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At",
           "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm"]
NUMBERS = {s: z for z, s in enumerate(SYMBOLS)}


class Structure:
    """positions [N,3] (Angstrom), cell [3,3] (rows = lattice vectors),
    numbers [N] (atomic numbers) or symbols, pbc [3]."""

    def __init__(self, positions, cell, numbers=None, symbols=None, pbc=True):
        self.positions = np.array(positions, dtype=np.float64).reshape(-1, 3)
        self.cell = np.array(cell, dtype=np.float64).reshape(3, 3)
        if numbers is None:
            if symbols is None:
                raise ValueError("numbers or symbols required")
            numbers = [NUMBERS[s] for s in symbols]
        self.numbers = np.array(numbers, dtype=np.int64).reshape(-1)
        if len(self.numbers) != len(self.positions):
            raise ValueError("numbers and positions differ in length")
        self.pbc = np.array([pbc] * 3 if np.isscalar(pbc) else pbc, dtype=bool).reshape(3)

    @property
    def symbols(self):
        return [SYMBOLS[z] for z in self.numbers]

    def get_chemical_symbols(self):
        return self.symbols

    def get_volume(self):
        return abs(float(np.linalg.det(self.cell)))

    def get_cell(self):
        return self.cell

    def get_positions(self):
        return self.positions

    def __len__(self):
        return len(self.positions)

    def copy(self):
        return Structure(self.positions.copy(), self.cell.copy(), self.numbers.copy(), pbc=self.pbc.copy())


def as_arrays(atoms):
    """(positions, cell, numbers, symbols, pbc) of any Atoms-like object."""
    pos = np.asarray(atoms.positions, dtype=np.float64).reshape(-1, 3)
    cell = np.asarray(atoms.cell, dtype=np.float64).reshape(3, 3)
    numbers = np.asarray(atoms.numbers, dtype=np.int64).reshape(-1)
    if hasattr(atoms, "get_chemical_symbols"):
        symbols = list(atoms.get_chemical_symbols())
    else:
        symbols = [SYMBOLS[z] for z in numbers]
    pbc = np.asarray(getattr(atoms, "pbc", True), dtype=bool)
    pbc = np.broadcast_to(pbc, (3,)).copy()
    return pos, cell, numbers, symbols, pbc
