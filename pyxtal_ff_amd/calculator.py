"""ASE-style calculator on the GPU adjoint path.

Mirrors the reference's ``PyXtalFFCalculator`` (``pyxtal_ff/calculator/
__init__.py:22-73``): energy, per-atom energies, forces and stress of one
structure from a descriptor and a per-element network, with the reference's
stress conventions -- ``stress_ml`` in GPa, Voigt order (xx, yy, zz, xy, xz,
yz) as ``calculate_properties`` returns it (``neuralnetwork.py:579-678``),
and for ``style='ase'`` ``stress = -(stress_ml * GPa)[[0, 1, 2, 5, 4, 3]]``
in eV/A^3.  The reference computes the descriptor, dxdr and rdxdr on the CPU
and contracts them with the network gradient; here the descriptor, its
adjoint (dE/dr per neighbour image) and the virial run on the GPU and dB/dR
is never formed.  No base (ZBL) potential.

Works with ASE ``Atoms`` when ASE is installed (it is then an
``ase.calculators.calculator.Calculator``) and with any object exposing
positions, cell, numbers and pbc otherwise.
"""
import numpy as np
import torch

from . import batch as bt
from . import potential as pot

EV2GPA = 160.21766208
GPA = 1.0 / EV2GPA  # ase.units.GPa in eV/A^3

try:  # pragma: no cover - ASE is optional
    from ase.calculators.calculator import Calculator as _Base, all_changes as _ALL
except ImportError:  # pragma: no cover
    _Base, _ALL = object, None


def make_descriptor(sb, descriptor, semantics="database"):
    """Batched descriptor of a StructureBatch from the reference's
    descriptor dict: {'type': 'SNAP'|'SO4'|'SO3', 'Rc', 'parameters':
    {'lmax'[, 'nmax', 'alpha', 'rfac', 'normalize_U']}, 'weights' (SNAP),
    'cutoff'}.

    The reference reads the dict differently in its two callers, and so
    does this function:
      * ``semantics="database"`` (``Database.compute``,
        ``utilities/__init__.py:186-220``): SNAP takes weights, lmax,
        normalize_U, cutoff and rfac; SO4 lmax, normalize_U and cutoff; SO3
        nmax, lmax and alpha (cutoff stays cosine);
      * ``semantics="calculator"`` (``compute_descriptor``, ``:276-308``,
        which the ASE calculator calls): the same, except that SNAP keeps
        the default rfac0 (0.99363) and SO3 the default alpha (2.0)."""
    if semantics not in ("database", "calculator"):
        raise ValueError("semantics must be 'database' or 'calculator'")
    t = descriptor["type"]
    rc = float(descriptor["Rc"])
    p = descriptor.get("parameters", {})
    cut = descriptor.get("cutoff", "cosine")
    if t in ("SNAP", "snap"):
        rfac0 = p.get("rfac", 0.99363) if semantics == "database" else 0.99363
        return bt.BatchedSNAP(sb, descriptor["weights"], lmax=p["lmax"], rcut=rc, rfac0=rfac0,
                              cutoff=cut, normalize_U=p.get("normalize_U", False))
    if t in ("SO4", "Bispectrum", "bispectrum"):
        return bt.BatchedSNAP(sb, None, lmax=p["lmax"], rcut=rc, rfac0=1.0, cutoff=cut,
                              normalize_U=p.get("normalize_U", False), site_weights="Z")
    if t in ("SO3", "SOAP", "soap"):
        alpha = p.get("alpha", 2.0) if semantics == "database" else 2.0
        return bt.BatchedSO3(sb, nmax=p["nmax"], lmax=p["lmax"], rcut=rc, alpha=alpha, cutoff="cosine")
    raise NotImplementedError(f"{t} is not implemented")


class PyXtalFFCalculator(_Base):
    implemented_properties = ["energy", "forces", "stress"]
    nolabel = True

    def __init__(self, descriptor, model, style="ase", device="cuda", **kwargs):
        """descriptor: the reference's descriptor dict (see make_descriptor);
        model: a ``potential.NNPotential`` (e.g. from
        ``lammps_io.potential_from_weights``)."""
        if _Base is not object:
            _Base.__init__(self, **kwargs)
        self.style = style
        self.descriptor = dict(descriptor)
        self.device = torch.device(device)
        self.model = model.to(self.device)
        self.results = {}
        self._key = None

    def _setup(self, atoms):
        """Structure metadata (cell, species, bin grid, descriptor state) is
        rebuilt only when the cell or the species change."""
        cell = np.asarray(atoms.cell, dtype=np.float64).reshape(3, 3)
        numbers = np.asarray(atoms.numbers)
        key = (cell.tobytes(), numbers.tobytes(), tuple(np.asarray(atoms.pbc, dtype=bool).reshape(-1)))
        if key != self._key:
            self._sb = bt.StructureBatch([atoms], float(self.descriptor["Rc"]), device=self.device)
            # the descriptor is built as the reference's compute_descriptor builds it
            self._desc = make_descriptor(self._sb, self.descriptor, semantics="calculator")
            self._species = pot.species_index(self._sb.symbols, self.model.elements, self.device)
            self._key = key
        else:
            pos = torch.as_tensor(np.asarray(atoms.positions, dtype=np.float64), device=self.device)
            self._sb.pos = pos
            self._sb.nl = self._sb.builder.build(pos)
            self._desc.state.nl = self._sb.nl

    def calculate(self, atoms=None, properties=("energy",), system_changes=_ALL):
        if _Base is not object:
            _Base.calculate(self, atoms, properties, system_changes)
            atoms = self.atoms
        self._setup(atoms)
        o = bt.energy_forces_virial(self._desc, self.model, self._species, virial=True)
        e_atom = o["atom_energies"]
        sv = bt.stress_voigt(o["virial"], self._sb)[0]  # eV/A^3, (xx, yy, zz, xy, xz, yz)
        stress_ml = (sv * EV2GPA).detach().cpu().numpy()
        self.results["energies"] = e_atom.detach().cpu().numpy()
        self.results["energy"] = float(o["energy"][0])
        self.results["free_energy"] = self.results["energy"]
        self.results["forces"] = o["forces"].detach().cpu().numpy()
        self.results["stress_ml"] = stress_ml
        self.results["energy_ml"] = self.results["energy"]
        self.results["forces_ml"] = self.results["forces"]
        if self.style == "ase":
            self.results["stress"] = -(stress_ml * GPA)[[0, 1, 2, 5, 4, 3]]
        else:
            self.results["stress"] = stress_ml

    # the ASE getters, for use without ASE
    def get_potential_energy(self, atoms=None, force_consistent=False):
        self.calculate(atoms, ["energy"])
        return self.results["energy"]

    def get_forces(self, atoms=None):
        self.calculate(atoms, ["forces"])
        return self.results["forces"]

    def get_stress(self, atoms=None):
        self.calculate(atoms, ["stress"])
        return self.results["stress"]

    def __str__(self):
        return "\nASE calculator with pyxtal_ff force field (gfx950 adjoint path)\n"

    __repr__ = __str__
