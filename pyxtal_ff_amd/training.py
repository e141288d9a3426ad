"""Force-matched training through the adjoint or on a materialised dxdr,
with the step captured in a HIP graph.

The loss is the reference's (``pyxtal_ff/models/neuralnetwork.py:466-545``,
``calculate_loss``), see ``reference_loss``.  The force term is
differentiated through the adjoint's double backward (``pxff_so4_tangent`` /
the SO3 tangent), so dB/dR is never materialised at any order.  Structures
are fixed during training, so the neighbour list, Z (or the SO3 ctot) and
the descriptors are computed once; each step runs the network, the adjoint
sweep (dE/dr per image, recomputed every step), its tangent, and the
optimiser.

With ``materialized=True`` the trainer instead does what the reference's
training does with its database records: dxdr (reference row layout) and
the per-atom rdxdr are computed once for the fixed structures and kept in
HBM, forces are the contraction of dE/dx with dxdr (``pxff_dxdr_contract``)
and its derivative with respect to dE/dx is the transposed contraction
(``pxff_dxdr_contract_t``); each step then streams dxdr twice instead of
re-running the adjoint sweep and its tangent.

Per-structure sums (energies, force errors, virials) are segment sums over
the contiguous atoms / images of each structure (``pxff_segment_sum``:
fixed order, no atomics, no memset nodes in a captured graph); only the
per-structure values (one per structure) go through torch reductions.
"""
import torch

from . import adjoint
from .batch import VOIGT

EV2GPA = 160.21766208  # neuralnetwork.py: eV2GPa


def reference_loss(E, E_ref, n_atoms, f_sq, sf, force_coefficient, S_gpa=None, S_ref=None, stress_mask=None,
                   stress_coefficient=None, reg=None):
    """The loss of ``calculate_loss`` (neuralnetwork.py:466-545) from
    per-structure quantities of one batch of S structures:

      E, E_ref [S]          predicted / label total energies (eV),
      n_atoms [S]           atoms per structure,
      f_sq [S]              sum over the structure's atoms of |F - F*|^2,
      sf [S]                per-structure weights (the softmax factors,
                            ones when no softmax is used),
      S_gpa, S_ref [S, 6]   predicted / label stress in GPa (Voigt xx yy zz
                            xy xz yz; the prediction is -sum dE/dx . rdxdr
                            times eV2GPa, :494-497, 506-507),
      stress_mask [S]       1 for structures whose group is in the stress
                            groups, else 0,
      reg                   alpha sum ||w||^2 over the weight matrices, or None.

    energy = sum sf ((E - E*) / n)^2 / (2 S); force = sum sf cF f_sq /
    (2 sum n); stress = sum mask sf cS |S - S*|^2 / (2 * 6 * sum mask).
    The terms are combined as the reference's branches combine them: the
    stress term only with a stress coefficient, the force term only with a
    force coefficient."""
    nst = E.shape[0]
    energy_loss = (sf * ((E - E_ref) / n_atoms) ** 2).sum() / (2.0 * nst)
    loss = energy_loss
    if force_coefficient:
        loss = loss + (sf * force_coefficient * f_sq).sum() / (2.0 * n_atoms.sum())
    if stress_coefficient:
        m = stress_mask if stress_mask is not None else torch.ones_like(E)
        s_count = 6.0 * m.sum()
        sl = (m * sf * stress_coefficient * ((S_gpa - S_ref) ** 2).sum(1)).sum()
        loss = loss + sl / (2.0 * torch.clamp(s_count, min=1.0))
    if reg is not None:
        loss = loss + reg
    return loss


def teacher_labels(desc, teacher, species, plain=True):
    """Energies [S] (eV), forces [N, 3] (eV/A) and Voigt stress [S, 6] (GPa,
    the unit of the reference's stress labels) of a teacher model,
    evaluated on the materialised (reference-layout) dxdr / rdxdr when
    ``plain`` (the reference's contraction, neuralnetwork.py:484-497), else
    through the adjoint."""
    from . import so4 as dso4, so3 as dso3
    b = desc.batch
    nl = b.nl.resolve()  # count settled before any kernel
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
        return (b.per_structure(e_atom.detach()), f.detach(),
                (b.per_structure(s_atom) * EV2GPA).detach())
    from .batch import energy_forces_virial, stress_voigt
    o = energy_forces_virial(desc, teacher, species, virial=True)
    return o["energy"].detach(), o["forces"].detach(), (stress_voigt(o["virial"], b) * EV2GPA).detach()


class _DxdrForce(torch.autograd.Function):
    """dE/dR [N, 3] = sum over rows (i, m) of beta_i . dxdr[(i, m)] (linear
    in beta; its backward is the transposed contraction)."""

    @staticmethod
    def forward(ctx, beta, trainer):
        from . import so4 as dso4
        ctx.trainer = trainer
        return dso4.dxdr_contract(trainer.batch.nl, trainer.nf, beta, trainer.dx)

    @staticmethod
    def backward(ctx, g):
        from . import so4 as dso4
        t = ctx.trainer
        return dso4.dxdr_contract_t(t.batch.nl, t.nf, g, t.dx), None


def weight_penalty(model, alpha):
    """alpha * sum of the squares of every parameter whose name contains
    'weight' (neuralnetwork.py:533-540), or None without alpha."""
    if not alpha:
        return None
    return alpha * sum((p * p).sum() for name, p in model.named_parameters() if "weight" in name)


class ForceTrainer:
    """Force-matched training of a per-element network on a fixed batch of
    structures (``NeuralNetwork.train`` with one batch of every structure).

    The loss is ``calculate_loss``'s (neuralnetwork.py:466-545, see
    ``reference_loss``): energies per atom over 2 x structures, forces over
    2 x atoms, stress in GPa over 2 x stress components, the per-structure
    weights ``sf`` and the alpha ||w||^2 penalty.  Energies come from the
    network on fixed descriptors, forces and the virial stress from the
    adjoint, or (``materialized``) from the dxdr and per-atom rdxdr computed
    once and kept in HBM.  labels = (E [S] eV, F [N, 3] eV/A, S [S, 6] GPa)."""

    def __init__(self, desc, model, species, labels, force_coefficient=0.1, stress_coefficient=0.01, alpha=None,
                 sf=None, stress_mask=None, lr=1e-3, graph=False, capturable=None, materialized=False):
        self.desc, self.model, self.species = desc, model, species
        b = desc.batch
        self.batch = b
        self.E_ref, self.F_ref, self.S_ref = labels
        self.cF, self.cS, self.alpha = force_coefficient, stress_coefficient, alpha
        dev = b.device
        self.sf = (torch.ones(b.nstruct, dtype=torch.float64, device=dev) if sf is None
                   else torch.as_tensor(sf, dtype=torch.float64, device=dev))
        # structures whose group is among the stress groups (all by default,
        # as get_stress_group collects every group, neuralnetwork.py:1032-1040)
        self.stress_mask = (torch.ones(b.nstruct, dtype=torch.float64, device=dev) if stress_mask is None
                            else torch.as_tensor(stress_mask, dtype=torch.float64, device=dev))
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
        self.materialized = bool(materialized)
        if self.materialized:
            from . import so4 as dso4
            with torch.no_grad():
                if hasattr(desc, "atom_w"):
                    self.nf = desc.cfg.nb
                    dx, rd = dso4.rows(nl, self.nf, dso4.dbdr(nl, desc.cfg, desc.atom_w, self.z), stress=True)
                else:
                    from . import so3 as dso3
                    self.nf = desc.cfg.ncoef
                    dx, rd = dso4.rows(nl, self.nf, dso3.dpdr(nl, desc.cfg, desc.state.atom_z, self.ct,
                                                              radial=self.rad), stress=True)
                self.dx = dx
                self.rv = dso4.rdxdr_reduce(nl, self.nf, rd)  # [N, nf, 6], Voigt
                del rd
        self.n_struct = torch.as_tensor(b.counts, dtype=torch.float64, device=dev)
        self._voigt = torch.as_tensor([3 * i + j for i, j in zip(*VOIGT)], dtype=torch.int64, device=dev)
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

    def predict(self):
        """Per-structure energies [S], forces [N, 3] and Voigt stress [S, 6]
        (GPa), differentiable in the network parameters."""
        b = self.batch
        nl = b.nl
        x = self.x.requires_grad_(True)
        e_atom = self.model.atom_energies(x, self.species)
        # per-structure buckets first: no multi-block sum over atoms (graph capture)
        (beta,) = torch.autograd.grad(b.per_structure(e_atom).sum(), [x], create_graph=True)
        if self.materialized:
            # neuralnetwork.py:484-497: F = -dE/dx . dxdr, S = -dE/dx . rdxdr (GPa)
            forces = -_DxdrForce.apply(beta.contiguous(), self)
            s_atom = -(beta.unsqueeze(2) * self.rv).sum(1)
            return b.per_structure(e_atom), forces, b.per_structure(s_atom) * EV2GPA
        dedr = self._dedr(beta.contiguous())
        forces = -adjoint._Gather.apply(dedr, nl)
        w = self.r.unsqueeze(2) * dedr.unsqueeze(1)  # r_e (x) dE/dr_e
        W = b.per_structure_images(w.reshape(-1, 9)).view(b.nstruct, 3, 3)
        # Voigt components by a device index (no host-to-device copy in a captured step)
        S = -W.view(b.nstruct, 9).index_select(1, self._voigt) * b.inv_vol_struct[:, None]
        return b.per_structure(e_atom), forces, S * EV2GPA

    def loss(self):
        """(loss, forces) of the current parameters."""
        E, forces, S = self.predict()
        f_sq = self.batch.per_structure(((forces - self.F_ref) ** 2).sum(1))
        loss = reference_loss(E, self.E_ref, self.n_struct, f_sq, self.sf, self.cF, S, self.S_ref,
                              self.stress_mask, self.cS, weight_penalty(self.model, self.alpha))
        return loss, forces

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
