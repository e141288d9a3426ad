"""LAMMPS ML-IAP export files (NN_weights.txt, DescriptorParam.txt): the
reference's shipped Si potentials read and written back byte for byte, and,
evaluated on the oracle's descriptors and dxdr, reproducing the DFT labels
they were trained on (a check of the file layout's meaning: layer order,
bias/weight layout, descriptor scaling)."""
import os

import numpy as np
import pytest
import torch

from pyxtal_ff_amd import lammps_io

HERE = os.path.dirname(os.path.abspath(__file__))
LMP = os.path.join(HERE, "golden", "lammps")
GOLD = os.path.join(HERE, "golden", "si_drange.npz")
MODELS = ("Si-snap-zbl", "Si-so3")


@pytest.mark.parametrize("model", MODELS)
def test_nn_weights_round_trip(tmp_path, model):
    src = os.path.join(LMP, model + "_NN_weights.txt")
    w = lammps_io.read_nn_weights(src)
    assert w["ndesc"] == 30 and w["layers"] == [("tanh", 12), ("tanh", 12), ("linear", 1)]
    assert w["nparams"] == 31 * 12 + 13 * 12 + 13
    pot = lammps_io.potential_from_weights(w)
    out = tmp_path / "NN_weights.txt"
    lammps_io.write_nn_weights(str(out), pot)
    assert out.read_text() == open(src).read()


@pytest.mark.parametrize("model", MODELS)
def test_descriptor_param_round_trip(tmp_path, model):
    src = os.path.join(LMP, model + "_DescriptorParam.txt")
    d = lammps_io.read_descriptor_param(src)
    if "twojmax" in d:
        info = {"type": "SNAP", "Rc": d["rcutfac"], "parameters": {"lmax": d["twojmax"] // 2, "rfac": d["rfac0"]},
                "weights": dict(zip(d["elems"], d["welems"]))}
        z = None
    else:
        info = {"type": "SO3", "Rc": d["rcutfac"], "parameters": {"nmax": d["nmax"], "lmax": d["lmax"],
                                                                  "alpha": d["alpha"]}}
        z = d["welems"]
    out = tmp_path / "DescriptorParam.txt"
    lammps_io.write_descriptor_param(str(out), info, d["elems"], atomic_numbers=z)
    assert out.read_text() == open(src).read()


def _structures(idx):
    g = np.load(GOLD)
    nat = g["natoms"]
    off = np.concatenate([[0], np.cumsum(nat)])
    for k in idx:
        yield (g["positions"][off[k]:off[k + 1]], g["cells"][k], g["pbc"][k], int(nat[k]), g["energy"][k],
               g["forces"][off[k]:off[k + 1]])


@pytest.mark.parametrize("model,idx", [("Si-snap-zbl", list(range(0, 239, 8))), ("Si-so3", [0, 50, 120, 213, 230])])
def test_shipped_potentials_reproduce_dft_labels(oracle_mod, model, idx):
    """Energies within 0.02 eV/atom (mean 0.008) and forces within a mean
    absolute error of 0.15 eV/A of the DFT labels (measured: 3.6 / 4.3 meV
    per atom, 0.09 / 0.08 eV/A for the SNAP / SO3 potential); a transposed
    layer or a wrong scaling is off by orders of magnitude."""
    from oracle import so3_oracle
    pot = lammps_io.potential_from_weights(lammps_io.read_nn_weights(os.path.join(LMP, model + "_NN_weights.txt")))
    de, df = [], []
    for pos, cell, pbc, n, e_dft, f_dft in _structures(idx):
        if "snap" in model:
            d = oracle_mod.snap_calculate(pos, cell, pbc, ["Si"] * n, {"Si": 1.0}, lmax=3, rcut=5.0, derivative=True)
        else:
            d = so3_oracle.calculate(pos, cell, pbc, np.full(n, 14), nmax=3, lmax=4, rcut=5.0, alpha=2.0)
        x = torch.as_tensor(d["x"]).requires_grad_(True)
        e = pot.atom_energies(x, [("Si", torch.arange(n))])
        (beta,) = torch.autograd.grad(e.sum(), [x])
        f = -np.einsum("ik,imkl->ml", beta.numpy(), oracle_mod.dense_dxdr(d))
        de.append(abs(float(e.sum().detach()) - e_dft) / n)
        df.append(np.abs(f - f_dft).mean())
    assert max(de) < 0.02 and np.mean(de) < 0.008, de
    assert np.mean(df) < 0.15, df
