"""The fused network kernel (pxff_mlp_energy_grad: per-element networks
forward and backward to dE/dx) against torch autograd on the same
parameters: three elements mixed over 300 atoms, narrow and deep / wide
networks (up to 8 layers of 64), a 147-long descriptor (SO3 nmax = lmax =
6: the matrix-core kernel with more than 64 KB of LDS), and atoms without a
network."""
import numpy as np
import pytest
import torch

from pyxtal_ff_amd import potential
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _torch_reference(model, x, symbols):
    xr = x.detach().clone().requires_grad_(True)
    sp = potential.species_index(symbols, model.elements, x.device)
    e = model.atom_energies(xr, sp)
    (g,) = torch.autograd.grad(e.sum(), [xr])
    return e.detach(), g


@pytest.mark.parametrize("hidden,act,nin", [((30, 30), "Tanh", 55), ((24, 16, 8), "Softplus", 55),
                                            ((64,) * 7, "ELU", 55), ((30, 30), "Tanh", 147)])
def test_fused_network_matches_autograd(hidden, act, nin):
    rng = np.random.default_rng(5)
    n = 300
    symbols = list(rng.choice(["O", "Si", "C"], size=n, p=[0.5, 0.4, 0.1]))
    x = torch.as_tensor(rng.random((n, nin)), dtype=torch.float64, device="cuda")
    dr = {e: np.stack([np.full(nin, -0.1), np.full(nin, 1.2)], 1) for e in ("O", "Si", "C")}
    model = potential.NNPotential(["O", "Si", "C"], nin, hidden, seed=7, drange=dr, activation=act).cuda()
    with torch.no_grad():
        for p in model.parameters():
            p.add_(0.05 * torch.randn_like(p))
    sp_atom = potential.species_atom(symbols, model.elements, "cuda")
    e, beta = model.energy_grad(x, sp_atom)
    e_ref, g_ref = _torch_reference(model, x, symbols)
    print("hidden", hidden, act, "rel err e", rel_err(e.cpu().numpy(), e_ref.cpu().numpy()),
          "beta", rel_err(beta.cpu().numpy(), g_ref.cpu().numpy()), flush=True)
    assert rel_err(e.cpu().numpy(), e_ref.cpu().numpy()) < 1e-13
    assert rel_err(beta.cpu().numpy(), g_ref.cpu().numpy()) < 1e-12


def test_fused_network_unknown_species_gives_nan():
    n, nin = 70, 10
    x = torch.rand((n, nin), dtype=torch.float64, device="cuda")
    model = potential.NNPotential(["Si"], nin, (8,), seed=1).cuda()
    sp = torch.zeros(n, dtype=torch.int64, device="cuda")
    sp[3] = 5
    e, beta = model.energy_grad(x, sp)
    assert torch.isnan(e[3]) and torch.isnan(beta[3]).all()
    keep = torch.ones(n, dtype=torch.bool, device="cuda")
    keep[3] = False
    assert torch.isfinite(e[keep]).all() and torch.isfinite(beta[keep]).all()
