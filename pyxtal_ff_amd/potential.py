"""Energy model on top of the descriptors: per-element MLP (FP64).

Mirrors the reference's per-element network of ``pyxtal_ff/models/
neuralnetwork.py`` (tanh hidden layers, linear output, descriptors scaled to
[0, 1] with the per-element drange, ``neuralnetwork.py:938-1029``) and its
energy / force / stress assembly (``calculate_properties``,
``neuralnetwork.py:579-678``), computed through either

  * the adjoint path: F = -dE/dR by ``torch.autograd`` through
    ``adjoint.bispectrum`` (never materialises dB/dR), or
  * the plain path: the reference's materialised dxdr [npairs, nb, 3] and the
    contraction F_m = -sum_rows dE/dx_i . dxdr[(i, m)].
"""
import numpy as np
import torch

from . import adjoint
from . import so4 as dev_so4

EV2GPA = 160.21766208


class ElementMLP(torch.nn.Module):
    def __init__(self, nin, hidden=(30, 30), seed=0):
        super().__init__()
        g = torch.Generator().manual_seed(int(seed))
        dims = [nin] + list(hidden) + [1]
        self.layers = torch.nn.ModuleList()
        for a, b in zip(dims[:-1], dims[1:]):
            lin = torch.nn.Linear(a, b, dtype=torch.float64)
            bound = float(np.sqrt(6.0 / (a + b)))  # Xavier uniform
            with torch.no_grad():
                lin.weight.copy_((torch.rand((b, a), generator=g, dtype=torch.float64) * 2 - 1) * bound)
                lin.bias.zero_()
            self.layers.append(lin)

    def forward(self, x):
        for k, lin in enumerate(self.layers):
            x = lin(x)
            if k < len(self.layers) - 1:
                x = torch.tanh(x)
        return x.squeeze(-1)


class NNPotential(torch.nn.Module):
    """Per-element MLPs; atoms of element e use models[e] on
    (x - dmin_e) / (dmax_e - dmin_e)."""

    def __init__(self, elements, nin, hidden=(30, 30), seed=0, drange=None):
        super().__init__()
        self.elements = list(elements)
        self.models = torch.nn.ModuleDict(
            {e: ElementMLP(nin, hidden, seed + 17 * k) for k, e in enumerate(self.elements)})
        for e in self.elements:
            lo = torch.zeros(nin, dtype=torch.float64)
            sc = torch.ones(nin, dtype=torch.float64)
            if drange is not None and e in drange:
                d = np.asarray(drange[e], dtype=np.float64)
                lo = torch.as_tensor(d[:, 0])
                sc = torch.as_tensor(1.0 / (d[:, 1] - d[:, 0]))
            self.register_buffer("dmin_" + e, lo)
            self.register_buffer("scale_" + e, sc)

    def atom_energies(self, x, species):
        """x [N, nb]; species: list of (element, index tensor)."""
        out = torch.zeros(x.shape[0], dtype=x.dtype, device=x.device)
        for e, idx in species:
            xe = x.index_select(0, idx)
            xe = (xe - getattr(self, "dmin_" + e)) * getattr(self, "scale_" + e)
            out = out.index_add(0, idx, self.models[e](xe))
        return out


def species_index(symbols, elements, device):
    sym = np.asarray(symbols)
    out = []
    for e in elements:
        idx = np.nonzero(sym == e)[0]
        if len(idx):
            out.append((e, torch.as_tensor(idx, dtype=torch.int64, device=device)))
    return out


def energy_forces_adjoint(nl, state, model, species, stress=False):
    """E (per structure sum), F [N,3], optional virial-based stress via the
    adjoint path."""
    pos = nl.pos.detach().requires_grad_(True)
    strain = torch.zeros((3, 3), dtype=torch.float64, device=pos.device, requires_grad=stress)
    r = adjoint.separations(nl, pos, strain if stress else None)
    b = adjoint.bispectrum(r, state)
    e_atom = model.atom_energies(b, species)
    e = e_atom.sum()
    inputs = [pos, strain] if stress else [pos]
    grads = torch.autograd.grad(e, inputs)
    f = -grads[0]
    out = {"energy": e.detach(), "forces": f, "atom_energies": e_atom.detach()}
    if stress:
        out["virial"] = grads[1]
    return out


def energy_forces_plain(nl, cfg, atom_w, site_w, model, species, stress=False):
    """Same quantities through the materialised dxdr path."""
    b, z, _ = dev_so4.forward(nl, cfg, atom_w, site_w)
    dbi = dev_so4.dbdr(nl, cfg, atom_w, z)
    dx, rd = dev_so4.rows(nl, cfg.nb, dbi, stress=stress)
    x = b.detach().requires_grad_(True)
    e_atom = model.atom_energies(x, species)
    e = e_atom.sum()
    (beta,) = torch.autograd.grad(e, [x])
    grad = dev_so4.dxdr_contract(nl, cfg.nb, beta, dx)
    out = {"energy": e.detach(), "forces": -grad, "atom_energies": e_atom.detach()}
    if stress:
        rv = dev_so4.rdxdr_reduce(nl, cfg.nb, rd)  # [N, nb, 6]
        out["stress_voigt"] = -torch.einsum("ik,ikl->l", beta, rv)
    return out
