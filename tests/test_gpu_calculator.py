"""The ASE-style calculator on the GPU adjoint path with the reference's
shipped Si potentials: energies, forces and stresses equal the oracle path
(descriptor, dxdr, per-atom rdxdr and the reference's contraction) to
1e-10, and reproduce the DFT labels the potentials were trained on."""
import os

import numpy as np
import pytest
import torch

from pyxtal_ff_amd import lammps_io
from pyxtal_ff_amd.calculator import EV2GPA, GPA, PyXtalFFCalculator
from pyxtal_ff_amd.structure import Structure
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LMP = os.path.join(HERE, "golden", "lammps")
GOLD = os.path.join(HERE, "golden", "si_drange.npz")
DESC = {"Si-snap-zbl": {"type": "SNAP", "Rc": 5.0, "parameters": {"lmax": 3}, "weights": {"Si": 1.0}},
        "Si-so3": {"type": "SO3", "Rc": 5.0, "parameters": {"nmax": 3, "lmax": 4, "alpha": 2.0}}}


@pytest.mark.parametrize("model", ["Si-snap-zbl", "Si-so3"])
def test_calculator_matches_oracle_and_labels(oracle_mod, model):
    from oracle import so3_oracle
    pot = lammps_io.potential_from_weights(lammps_io.read_nn_weights(os.path.join(LMP, model + "_NN_weights.txt")))
    calc = PyXtalFFCalculator(DESC[model], pot)
    g = np.load(GOLD)
    nat = g["natoms"]
    off = np.concatenate([[0], np.cumsum(nat)])
    for k in (3, 77, 150, 225):
        n = int(nat[k])
        s = Structure(g["positions"][off[k]:off[k + 1]], g["cells"][k], numbers=[14] * n, pbc=g["pbc"][k])
        e = calc.get_potential_energy(s)
        f = calc.get_forces(s)
        st = calc.get_stress(s)
        sml = calc.results["stress_ml"]
        # the oracle path: reference descriptor arrays and the reference's contraction
        if "snap" in model:
            d = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, ["Si"] * n, {"Si": 1.0}, lmax=3, rcut=5.0,
                                          derivative=True, stress=True)
        else:
            d = so3_oracle.calculate(s.positions, s.cell, s.pbc, np.full(n, 14), nmax=3, lmax=4, rcut=5.0,
                                     alpha=2.0, derivative=True, stress=True)
        pc = pot.cpu()
        x = torch.as_tensor(d["x"]).requires_grad_(True)
        ea = pc.atom_energies(x, [("Si", torch.arange(n))])
        (beta,) = torch.autograd.grad(ea.sum(), [x])
        f_ref = -np.einsum("ik,imkl->ml", beta.numpy(), oracle_mod.dense_dxdr(d))
        s_ref = -np.einsum("ik,ikl->l", beta.numpy(), oracle_mod.rdxdr_per_atom(d)) * EV2GPA
        pot.to("cuda")
        assert abs(e - float(ea.sum().detach())) < 1e-10 * abs(e)
        assert rel_err(f, f_ref) < 1e-9
        assert rel_err(sml, s_ref) < 1e-9
        assert np.allclose(st, -(sml * GPA)[[0, 1, 2, 5, 4, 3]], rtol=0, atol=1e-15)
        # the labels the shipped potential was fitted to (the SNAP model
        # was fitted on top of a ZBL core repulsion, which is not added
        # here: close-contact atoms carry most of the force error)
        assert abs(e - g["energy"][k]) / n < 0.02
        assert np.abs(f - g["forces"][off[k]:off[k + 1]]).mean() < 0.3


def test_calculator_reuses_metadata_and_tracks_positions():
    """A second call with moved atoms (same cell) reuses the structure set-up
    and gives the forces of the new positions (central difference of E)."""
    pot = lammps_io.potential_from_weights(lammps_io.read_nn_weights(os.path.join(LMP, "Si-snap-zbl_NN_weights.txt")))
    calc = PyXtalFFCalculator(DESC["Si-snap-zbl"], pot)
    g = np.load(GOLD)
    n = int(g["natoms"][0])
    s = Structure(g["positions"][:n], g["cells"][0], numbers=[14] * n, pbc=g["pbc"][0])
    f = calc.get_forces(s).copy()
    key = calc._key
    h = 1e-5
    for a, c in ((5, 0), (40, 2)):
        es = []
        for sg in (1, -1):
            s2 = s.copy()
            s2.positions[a, c] += sg * h
            es.append(calc.get_potential_energy(s2))
        assert calc._key == key
        assert abs(-(es[0] - es[1]) / (2 * h) - f[a, c]) < 1e-6 * max(1.0, abs(f[a, c]))
