"""Force-matched training through the adjoint, with the step captured in a
HIP graph.

Mirrors the reference's loss (``pyxtal_ff/models/neuralnetwork.py:466-545``):
energy, force and stress terms per atom with user coefficients.  The force
term is differentiated through the adjoint's double backward
(``pxff_so4_tangent`` / the SO3 tangent), so dB/dR is never materialised at
any order.  Structures are fixed during training, so the neighbour list, Z
(or the SO3 ctot) and the descriptors are computed once; each step runs the
network, the adjoint sweep (dE/dr per image, recomputed every step), its
tangent, and the optimiser.

Per-structure sums (energies, force errors, virials) are segment sums over
the contiguous atoms / images of each structure (``pxff_segment_sum``:
fixed order, no atomics, no memset nodes in a captured graph); only the
per-structure values (one per structure) go through torch reductions.
"""
import torch

from . import adjoint
from .batch import VOIGT


def teacher_labels(desc, teacher, species, plain=True):
    """Energies [S], forces [N, 3], Voigt stress [S, 6] of a teacher model,
    evaluated on the materialised (reference-layout) dxdr / rdxdr when
    ``plain`` (the reference's contraction, neuralnetwork.py:484-497), else
    through the adjoint."""
    from . import so4 as dso4, so3 as dso3
    b = desc.batch
    nl = b.nl
    if plain and hasattr(desc, "atom_w"):
        x, z, _ = dso4.forward(nl, desc.cfg, desc.atom_w, desc.site_w)
        dx, rd = dso4.rows(nl, desc.cfg.nb, dso4.dbdr(nl, desc.cfg, desc.atom_w, z), stress=True)
        nf = desc.cfg.nb
    elif plain:
        x, ct, rad = dso3.forward(nl, desc.cfg, desc.state.atom_z, radial=True)
        dx, rd = dso4.rows(nl, desc.cfg.ncoef, dso3.dpdr(nl, desc.cfg, desc.state.atom_z, ct, radial=rad),
                           stress=True)
        nf = desc.cfg.ncoef
    if plain:
        xx = x.detach().requires_grad_(True)
        e_atom = teacher.atom_energies(xx, species)
        (beta,) = torch.autograd.grad(e_atom.sum(), [xx])
        f = -dso4.dxdr_contract(nl, nf, beta, dx)
        rv = dso4.rdxdr_reduce(nl, nf, rd)  # [N, nf, 6]
        s_atom = -torch.einsum("ik,ikl->il", beta, rv)
        return b.per_structure(e_atom.detach()), f.detach(), b.per_structure(s_atom).detach()
    from .batch import energy_forces_virial, stress_voigt
    o = energy_forces_virial(desc, teacher, species, virial=True)
    return o["energy"].detach(), o["forces"].detach(), stress_voigt(o["virial"], b).detach()


class ForceTrainer:
    """Loss = mean over structures of
         cE ((E - E*) / n)^2 + cF sum_i |F_i - F*_i|^2 / (3 n) + cS |S - S*|^2 / 6
    (per atom), with E from the network on fixed descriptors, F and the
    virial stress from the adjoint."""

    def __init__(self, desc, model, species, labels, coef=(1.0, 0.1, 0.01), lr=1e-3, graph=False,
                 capturable=None):
        self.desc, self.model, self.species = desc, model, species
        b = desc.batch
        self.batch = b
        self.E_ref, self.F_ref, self.S_ref = labels
        self.cE, self.cF, self.cS = coef
        nl = b.nl
        self.r = nl.img_r[: 3 * nl.nimg].view(nl.nimg, 3).clone()
        # fixed structures: descriptors and (for SO4) Z are computed once
        with torch.no_grad():
            self.x = desc.descriptors(self.r).detach()
        if hasattr(desc, "atom_w"):
            from . import so4 as dso4
            _, self.z, _ = dso4.forward(nl, desc.cfg, desc.atom_w, desc.site_w, desc.batch.work_order)
        else:
            from . import so3 as dso3
            # fixed structures: the radial integrals are computed once too
            _, self.ct, self.rad = dso3.forward(nl, desc.cfg, desc.state.atom_z, radial=True)
        self.n_struct = torch.as_tensor(b.counts, dtype=torch.float64, device=b.device)
        self._voigt = torch.as_tensor([3 * i + j for i, j in zip(*VOIGT)], dtype=torch.int64, device=b.device)
        # Adam keeps its step count on the device when the step is captured;
        # its update then differs from the host-count variant in the last
        # bits, so eager runs meant to match graph replays pass capturable=True
        self.opt = torch.optim.Adam(model.parameters(), lr=lr, capturable=graph if capturable is None else capturable)
        self.graph = None
        self.use_graph = graph
        self.static_loss = None

    def _dedr(self, beta):
        if hasattr(self, "z"):
            return adjoint._DEdR.apply(beta, self.r, self.z, self.desc.state)
        from .so3 import _SO3DEdR
        return _SO3DEdR.apply(beta, self.r, self.ct, self.desc.state, self.rad)

    def loss(self):
        b = self.batch
        nl = b.nl
        x = self.x.requires_grad_(True)
        e_atom = self.model.atom_energies(x, self.species)
        # per-structure buckets first: no multi-block sum over atoms (graph capture)
        (beta,) = torch.autograd.grad(b.per_structure(e_atom).sum(), [x], create_graph=True)
        dedr = self._dedr(beta.contiguous())
        forces = -adjoint._Gather.apply(dedr, nl)
        w = self.r.unsqueeze(2) * dedr.unsqueeze(1)  # r_e (x) dE/dr_e
        W = b.per_structure_images(w.reshape(-1, 9)).view(b.nstruct, 3, 3)
        # Voigt components by a device index (no host-to-device copy in a captured step)
        S = -W.view(b.nstruct, 9).index_select(1, self._voigt) * b.inv_vol_struct[:, None]
        E = b.per_structure(e_atom)
        le = ((E - self.E_ref) / self.n_struct) ** 2
        fa = ((forces - self.F_ref) ** 2).sum(1)
        lf = b.per_structure(fa) / (3 * self.n_struct)
        ls = ((S - self.S_ref) ** 2).sum(1) / 6
        per_s = self.cE * le + self.cF * lf + self.cS * ls
        return per_s.mean(), forces

    def _step_eager(self):
        self.opt.zero_grad(set_to_none=False)
        loss, _ = self.loss()
        loss.backward()
        self.opt.step()
        return loss

    def step(self):
        if not self.use_graph:
            return self._step_eager().detach()
        if self.graph is None:
            # warm up on a side stream, then capture one full step
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                for _ in range(2):
                    self._step_eager()
            torch.cuda.current_stream().wait_stream(s)
            self.graph = torch.cuda.CUDAGraph()
            self.opt.zero_grad(set_to_none=False)
            with torch.cuda.graph(self.graph):
                self.static_loss = self._step_eager()
        self.graph.replay()
        return self.static_loss.detach()
