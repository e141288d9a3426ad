"""Energy model on top of the descriptors: per-element MLP (FP64).

Mirrors the reference's per-element network of ``pyxtal_ff/models/
neuralnetwork.py`` (tanh hidden layers, linear output, descriptors scaled to
[0, 1] with the per-element drange, ``neuralnetwork.py:938-1029``) and its
energy / force / stress assembly (``calculate_properties``,
``neuralnetwork.py:579-678``), computed through either

  * the adjoint path: F = -dE/dR by ``torch.autograd`` through
    ``adjoint.bispectrum`` (never materialises dB/dR; differentiable in the
    network parameters, for training), or
  * the fused adjoint path: the same kernels called directly, with the
    networks evaluated forward and backward to dE/dB by one kernel
    (``pxff_mlp_energy_grad``) -- energy + force evaluation, or
  * the plain path: the reference's materialised dxdr [npairs, nb, 3] and the
    contraction F_m = -sum_rows dE/dx_i . dxdr[(i, m)].
"""
import numpy as np
import torch

from . import adjoint
from . import so3 as dev_so3
from . import so4 as dev_so4

EV2GPA = 160.21766208


# activations of the reference's networks (neuralnetwork.py:105-109) that
# the fused kernel (pxff_mlp_energy_grad) evaluates, with its ids
ACTIVATIONS = {"Linear": 0, "Tanh": 1, "Sigmoid": 2, "ReLU": 3, "Softplus": 4, "ELU": 5}
_TORCH_ACT = {"Linear": lambda v: v, "Tanh": torch.tanh, "Sigmoid": torch.sigmoid,
              "ReLU": torch.relu, "Softplus": torch.nn.functional.softplus, "ELU": torch.nn.functional.elu}


_MLP_SPLIT, _MLP_SPLIT_ROWS = 64, 8192


class ElementMLP(torch.nn.Module):
    """nn.Linear layers, one activation after each hidden layer, a linear
    output layer of width 1 (neuralnetwork.py:186-200)."""

    def __init__(self, nin, hidden=(30, 30), seed=0, activation="Tanh", out_activation="Linear"):
        super().__init__()
        g = torch.Generator().manual_seed(int(seed))
        dims = [nin] + list(hidden) + [1]
        acts = [activation] * len(hidden) if isinstance(activation, str) else list(activation)
        if len(acts) != len(hidden):
            raise ValueError("one activation per hidden layer")
        for a in acts + [out_activation]:
            if a not in ACTIVATIONS:
                raise NotImplementedError(f"activation {a} (available: {sorted(ACTIVATIONS)})")
        self.acts = acts + [out_activation]
        self.layers = torch.nn.ModuleList()
        for a, b in zip(dims[:-1], dims[1:]):
            lin = torch.nn.Linear(a, b, dtype=torch.float64)
            bound = float(np.sqrt(6.0 / (a + b)))  # Xavier uniform
            with torch.no_grad():
                lin.weight.copy_((torch.rand((b, a), generator=g, dtype=torch.float64) * 2 - 1) * bound)
                lin.bias.zero_()
            self.layers.append(lin)

    def forward(self, x):
        n = x.shape[0]
        if x.dim() == 2 and n >= _MLP_SPLIT_ROWS:
            # the rows as _MLP_SPLIT batches of one batched product: the
            # weight gradients are then _MLP_SPLIT parallel products summed
            # afterwards, not one product over every atom (which one
            # workgroup walked through, ~2.4 ms per layer at 34k atoms)
            c = _MLP_SPLIT
            rows = -(-n // c)
            h = torch.nn.functional.pad(x, (0, 0, 0, c * rows - n)).view(c, rows, x.shape[1])
            for lin, act in zip(self.layers, self.acts):
                wt = lin.weight.t().unsqueeze(0).expand(c, -1, -1).contiguous()
                h = _TORCH_ACT[act](torch.bmm(h, wt) + lin.bias)
            return h.reshape(c * rows, -1)[:n].squeeze(-1)
        for lin, act in zip(self.layers, self.acts):
            x = _TORCH_ACT[act](lin(x))
        return x.squeeze(-1)


class NNPotential(torch.nn.Module):
    """Per-element MLPs; atoms of element e use models[e] on
    (x - dmin_e) / (dmax_e - dmin_e)."""

    def __init__(self, elements, nin, hidden=(30, 30), seed=0, drange=None, activation="Tanh",
                 out_activation="Linear"):
        super().__init__()
        self.elements = list(elements)
        self.nin = int(nin)
        self.models = torch.nn.ModuleDict(
            {e: ElementMLP(nin, hidden, seed + 17 * k, activation, out_activation)
             for k, e in enumerate(self.elements)})
        self._packed = None
        for e in self.elements:
            lo = torch.zeros(nin, dtype=torch.float64)
            rng = torch.ones(nin, dtype=torch.float64)
            if drange is not None and e in drange:
                d = np.asarray(drange[e], dtype=np.float64)
                lo = torch.as_tensor(d[:, 0])
                rng = torch.as_tensor(d[:, 1] - d[:, 0])
            self.register_buffer("dmin_" + e, lo)
            self.register_buffer("range_" + e, rng)  # dmax - dmin (Scale1 of the LAMMPS export)
            self.register_buffer("scale_" + e, 1.0 / rng)

    def atom_energies(self, x, species):
        """x [N, nb]; species: list of (element, index tensor)."""
        out = torch.zeros(x.shape[0], dtype=x.dtype, device=x.device)
        for e, idx in species:
            xe = x.index_select(0, idx)
            xe = (xe - getattr(self, "dmin_" + e)) * getattr(self, "scale_" + e)
            out = out.index_add(0, idx, self.models[e](xe))
        return out

    def packed(self):
        """Parameters in the fused kernel's layout: (host layout int32,
        device layout, params, dmin [nelem, nin], scale [nelem, nin]).
        Rebuilt when any parameter or buffer changed in place (optimiser
        steps bump tensor versions)."""
        # the module tree is walked once; the key then costs one pair per
        # tensor (walking it on every call took most of the host time of an
        # E+F step)
        ts = self.__dict__.get("_packed_tensors")
        if ts is None:
            ts = list(self.parameters()) + list(self.buffers())
            self.__dict__["_packed_tensors"] = ts
        key = tuple((t.data_ptr(), t._version) for t in ts)
        if self._packed is not None and self._packed[0] == key:
            return self._packed[1]
        stride = 1 + 6 * 8
        lay = np.zeros((len(self.elements), stride), dtype=np.int32)
        chunks, off = [], 0
        for ei, e in enumerate(self.elements):
            m = self.models[e]
            if len(m.layers) > 8:
                raise NotImplementedError("the fused network kernel takes at most 8 layers")
            lay[ei, 0] = len(m.layers)
            for k, (lin, act) in enumerate(zip(m.layers, m.acts)):
                w = lin.weight.detach().reshape(-1)
                b = lin.bias.detach().reshape(-1)
                wt = lin.weight.detach().t().contiguous().reshape(-1)
                lay[ei, 1 + 6 * k: 7 + 6 * k] = [lin.in_features, lin.out_features, ACTIVATIONS[act], off,
                                                 off + w.numel(), off + w.numel() + b.numel()]
                chunks += [w, b, wt]
                off += 2 * w.numel() + b.numel()
        params = torch.cat(chunks).contiguous()
        dev = params.device
        dmin = torch.stack([getattr(self, "dmin_" + e) for e in self.elements]).to(dev).contiguous()
        scale = torch.stack([getattr(self, "scale_" + e) for e in self.elements]).to(dev).contiguous()
        out = (lay, torch.as_tensor(lay, device=dev), params, dmin, scale)
        self._packed = (key, out)
        return out

    def energy_grad(self, x, species_atom):
        """Atomic energies [N] and dE/dx [N, nin] by the fused kernel (no
        autograd; parameters are constants).  species_atom: int64 [N],
        index into ``elements``."""
        from . import _native as nat
        lay, dlay, params, dmin, scale = self.packed()
        n = x.shape[0]
        x = x.contiguous()
        e = torch.empty(max(n, 1), dtype=torch.float64, device=x.device)
        beta = torch.empty((max(n, 1), self.nin), dtype=torch.float64, device=x.device)
        nat.check(nat.lib().pxff_mlp_energy_grad(
            n, self.nin, len(self.elements), nat.ptr(species_atom), lay.ctypes.data_as(nat.P), nat.ptr(dlay),
            nat.ptr(params), nat.ptr(dmin), nat.ptr(scale), nat.ptr(x), nat.ptr(e), nat.ptr(beta), nat.stream()),
            "pxff_mlp_energy_grad")
        return e[:n], beta[:n]


def species_atom(symbols, elements, device):
    """Per-atom index into ``elements`` (int64 device tensor)."""
    pos = {e: k for k, e in enumerate(elements)}
    return torch.as_tensor(np.array([pos[s] for s in symbols], dtype=np.int64), device=device)


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


def _fused_adjoint(nl, cfg, weights, model, species_atom, work_order=None, counted=False):
    """The descriptor kernels and the fused network of one E+F step up to
    the per-image sweep: (atomic energies, a function launching the sweep
    that returns dE/dr per image).  weights: (atom_w, site_w) for SO(4) /
    SNAP, (atomic numbers,) for SO(3)."""
    if isinstance(cfg, dev_so3.So3Config):
        (atom_z,) = weights
        x, ct, rad = dev_so3.forward(nl, cfg, atom_z, radial=True)
        e_atom, beta = model.energy_grad(x, species_atom)
        a = dev_so3.adjoint_a(nl, cfg, ct, beta)
        return e_atom, lambda: dev_so3.deidrj(nl, cfg, atom_z, a, radial=rad, counted=counted)
    atom_w, site_w = weights
    b, z, _ = dev_so4.forward(nl, cfg, atom_w, site_w, work_order)
    e_atom, beta = model.energy_grad(b, species_atom)
    y = dev_so4.yi(nl, cfg, z, beta)
    return e_atom, lambda: dev_so4.deidrj(nl, cfg, atom_w, y, counted=counted)


def state_weights(state):
    """The weight tensors of an SO4State ((atom_w, site_w)) or SO3State ((atom_z,))."""
    return (state.atom_z,) if isinstance(state, dev_so3.SO3State) else (state.atom_w, state.site_w)


def energy_forces_fused(nl, state, model, species_atom, virial=False):
    """E, F (and the per-centre virial sum) through the adjoint kernels with
    no autograd: descriptor (SO(4): U_tot -> Z -> B; SO(3): expansion ->
    power spectrum) -> network energies and dE/dx (fused kernel) -> Y or A
    -> dE/dr per image -> forces by the deterministic gather.  Same
    quantities as ``energy_forces_adjoint``; the network parameters are
    constants here (energy + force evaluation, e.g. MD or the calculator).
    state: an ``adjoint.SO4State`` or a ``so3.SO3State``."""
    gen = nl.generation
    e_atom, sweep = _fused_adjoint(nl, state.cfg, state_weights(state), model, species_atom,
                                   getattr(state, "work_order", None))
    # the first use of the host-side image count (the build did not wait
    # for it); if the list had to be rebuilt, everything above used a cut list
    nl.nimg
    if nl.generation != gen:
        return energy_forces_fused(nl, state, model, species_atom, virial)
    dedr = sweep()
    # forces and the energy total (one extra workgroup) in one launch
    f, energy, vir = dev_so4.forces_energy(nl, dedr, e_atom, virial=virial)
    out = {"energy": energy[0], "atom_energies": e_atom, "forces": f}
    if virial:
        out["virial"] = vir.sum(0)
    return out


def energy_forces_plain(nl, cfg, atom_w, site_w, model, species, stress=False):
    """Same quantities through the materialised dxdr path."""
    nl.resolve()  # count settled before any kernel
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


def _segment_total(v, seg):
    """Column sums of v [rows, ncol] over the one segment seg = [0, rows]
    (pxff_segment_sum: fixed reduction order, no atomics, no memset)."""
    from . import _native as nat
    vv = v.contiguous()
    ncol = vv.shape[1] if vv.dim() > 1 else 1
    out = torch.empty(ncol, dtype=torch.float64, device=v.device)
    nat.check(nat.lib().pxff_segment_sum(1, ncol, nat.ptr(seg), nat.ptr(vv), nat.ptr(out), nat.stream()),
              "pxff_segment_sum")
    return out


class FusedEFGraph:
    """Energy, forces (and the virial) of a fixed set of atoms as one HIP
    graph replay: the neighbour-list rebuild from the positions in HBM and
    the fused adjoint kernels of ``energy_forces_fused`` (SO(4)/SNAP or
    SO(3)), captured once and replayed per step, so no host work sits
    between the kernels and the step never waits for the host.

    cfg: an ``So4Config`` with weights = (atom_w, site_w), or an
    ``So3Config`` with weights = (atomic numbers,).

    The captured list has a fixed image capacity (the builder's).  Each
    replay's unclamped image count is copied back asynchronously; with
    ``check`` (default) the call waits for it, and if it exceeded the
    capacity the capacity is grown, the step re-captured and replayed (the
    recovery ``DeviceNeighborList.nimg`` does for eager builds).  The
    returned tensors belong to the graph: the next replay overwrites them.
    Per-image work is launched over the capacity and bounded by the image
    count on the device (``pxff_so4_deidrj_counted`` /
    ``pxff_so3_deidrj_counted``)."""

    def __init__(self, builder, cfg, weights, model, species_atom, pos, virial=False, warmup=2):
        self.builder, self.cfg, self.weights = builder, cfg, tuple(weights)
        self.model, self.species_atom, self.virial = model, species_atom, virial
        self.pos = pos.detach().clone().contiguous()
        self._seg = torch.tensor([0, builder.natoms], dtype=torch.int64, device=pos.device)
        self._count_host = torch.empty(1, dtype=torch.int64, pin_memory=True)
        self._ev = torch.cuda.Event()
        self.captures = 0
        self._capture(warmup)

    def _body(self):
        n, cfg = self.builder.natoms, self.cfg
        nl = self.builder.build(self.pos, host_count=False)
        e_atom, sweep = _fused_adjoint(nl, cfg, self.weights, self.model, self.species_atom, counted=True)
        dedr = sweep()
        f, energy, vir = dev_so4.forces_energy(nl, dedr, e_atom, virial=self.virial)
        out = {"atom_energies": e_atom, "energy": energy[0], "forces": f}
        if self.virial:
            out["virial"] = _segment_total(vir.reshape(n, 9), self._seg).view(3, 3)
        self._nl = nl
        return out

    def _capture(self, warmup):
        for _ in range(warmup):  # tables, tuned launch parameters, allocator pools
            self._body()
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = self._body()
        self.captures += 1

    def __call__(self, pos=None, check=True):
        if pos is not None:
            self.pos.copy_(pos)
        self.graph.replay()
        self._count_host.copy_(self._nl.total_device, non_blocking=True)
        self._ev.record()
        if check:
            self.check()
        return self.out

    def check(self):
        """Wait for the last replay's image count; redo the step with a
        larger capacity if the list was cut.  Returns the image count."""
        while True:
            self._ev.synchronize()
            total = int(self._count_host[0])
            if total <= self._nl.img_capacity:
                return total
            b = self.builder
            b.img_capacity = max(b.img_capacity, total + total // 4 + 64)
            self._capture(1)
            self.graph.replay()
            self._count_host.copy_(self._nl.total_device, non_blocking=True)
            self._ev.record()
