"""Batched descriptors, energies, forces and virials of many structures.

The reference computes one structure per call and fans structures out over a
``multiprocessing.Pool`` (``pyxtal_ff/utilities/__init__.py:121-164``).  Here
a batch of structures of uneven atom and neighbour counts is one image list
(``NeighborListBuilder`` over all cells), and each kernel runs once over the
whole batch: the per-centre U_tot kernel walks a neighbour-count-sorted work
list persistently (heaviest centres first), the Z kernel walks 64-atom tiles
of the concatenated atoms, and the image kernels are flat over all images.
"""
import numpy as np
import torch

from . import _native as nat
from . import adjoint
from . import so3 as dev_so3
from . import so4 as dev_so4
from .neighbors import NeighborListBuilder
from .structure import as_arrays


class SegmentSum(torch.autograd.Function):
    """out[s] = sum of rows [seg[s], seg[s+1]) of v (pxff_segment_sum: one
    workgroup per segment, fixed reduction order, no atomics); backward
    broadcasts the segment gradient back to its rows."""

    @staticmethod
    def forward(ctx, v, seg, row_seg):
        ctx.row_seg = row_seg
        nseg = seg.numel() - 1
        vv = v.contiguous()
        ncol = int(np.prod(v.shape[1:])) if v.dim() > 1 else 1
        out = torch.empty((nseg,) + tuple(v.shape[1:]), dtype=v.dtype, device=v.device)
        nat.check(nat.lib().pxff_segment_sum(nseg, ncol, nat.ptr(seg), nat.ptr(vv), nat.ptr(out), nat.stream()),
                  "pxff_segment_sum")
        return out

    @staticmethod
    def backward(ctx, g):
        return g.index_select(0, ctx.row_seg), None, None


def segment_sum(v, seg, row_seg):
    """Sum of the rows of v over contiguous segments (differentiable)."""
    if v.dtype != torch.float64:
        raise TypeError("segment_sum expects float64")
    return SegmentSum.apply(v, seg, row_seg)


class StructureBatch:
    """Concatenated atoms of a list of Atoms-like structures on the GPU."""

    def __init__(self, structures, rc, device="cuda"):
        arrs = [as_arrays(s) for s in structures]
        self.counts = np.array([len(a[0]) for a in arrs], dtype=np.int64)
        self.nstruct = len(arrs)
        self.symbols = [sym for a in arrs for sym in a[3]]
        self.numbers = np.concatenate([a[2] for a in arrs]) if arrs else np.zeros(0, np.int64)
        cells = np.stack([a[1] for a in arrs])
        pbcs = np.stack([a[4] for a in arrs])
        self.volumes = np.abs(np.linalg.det(cells))
        dev = torch.device(device)
        self.device = dev
        self.pos = torch.as_tensor(np.concatenate([a[0] for a in arrs]), dtype=torch.float64, device=dev)
        self.builder = NeighborListBuilder(self.counts, cells, pbcs, rc, device)
        self.nl = self.builder.build(self.pos).resolve()  # counts final before the work list is sorted
        nl = self.nl
        self.atom_struct = self.builder.atom_struct
        self.natoms = nl.natoms
        # neighbour-count-sorted work list (heaviest centres first)
        k = nl.img_start[1:] - nl.img_start[:-1]
        self.work_order = torch.argsort(k, descending=True, stable=True)
        self.atom_seg = self.builder.atom_start
        self.inv_vol_struct = torch.as_tensor(1.0 / self.volumes, dtype=torch.float64, device=dev)
        self._img_cache = None

    def _img_segments(self):
        """(image segment starts [nstruct+1], structure of each image) of the
        current list; images are grouped by centre, centres by structure."""
        nl = self.nl
        if self._img_cache is None or self._img_cache[0] is not nl:
            seg = nl.img_start.index_select(0, self.atom_seg)
            st = self.atom_struct.index_select(0, nl.img_atom[: nl.nimg])
            self._img_cache = (nl, seg, st)
        return self._img_cache[1], self._img_cache[2]

    @property
    def img_struct(self):
        return self._img_segments()[1]

    def per_structure(self, v):
        """Sum of a per-atom tensor [N, ...] over each structure -> [nstruct, ...]."""
        return segment_sum(v, self.atom_seg, self.atom_struct)

    def per_structure_images(self, v):
        """Sum of a per-image tensor [nimg, ...] over each structure."""
        seg, st = self._img_segments()
        return segment_sum(v, seg, st)


class BatchedSNAP:
    """SNAP bispectrum (SNAP.py conventions) of a StructureBatch."""

    def __init__(self, batch, weights, lmax=4, rcut=4.9, rfac0=0.99363, cutoff="cosine", normalize_U=False,
                 store_z=True, site_weights=None):
        """weights: per-element neighbour weight (SNAP.py), or None for the
        atomic numbers (SO4.py); site_weights: per-element self weight, 1.0
        by default (SNAP.py), or "Z" for the atomic numbers (SO4.py)."""
        self.batch = batch
        self.cfg = dev_so4.So4Config(int(round(2 * lmax)), rcut, rfac0, cutoff, normalize_U)
        dev = batch.device
        z = np.asarray(batch.numbers, dtype=np.float64)
        aw = z if weights is None else np.array([weights[s] for s in batch.symbols], dtype=np.float64)
        sw = z if site_weights == "Z" else np.ones(batch.natoms)
        self.atom_w = torch.as_tensor(aw, dtype=torch.float64, device=dev)
        self.site_w = torch.as_tensor(sw, dtype=torch.float64, device=dev)
        self.state = adjoint.SO4State(batch.nl, self.cfg, self.atom_w, self.site_w, store_z=store_z,
                                      work_order=batch.work_order)

    @property
    def nfeat(self):
        return self.cfg.nb

    def descriptors(self, r):
        return adjoint.bispectrum(r, self.state)


class BatchedSO3:
    """SO3 power spectrum of a StructureBatch."""

    def __init__(self, batch, nmax=6, lmax=6, rcut=5.0, alpha=2.0, cutoff="cosine", weight_on=False):
        self.batch = batch
        self.cfg = dev_so3.So3Config(nmax, lmax, rcut, alpha, cutoff, weight_on)
        z = torch.as_tensor(batch.numbers, dtype=torch.int64, device=batch.device)
        self.state = dev_so3.SO3State(batch.nl, self.cfg, z)

    @property
    def nfeat(self):
        return self.cfg.ncoef

    def descriptors(self, r):
        return dev_so3.power_spectrum(r, self.state)


def energy_forces_virial(desc, model, species, virial=True, create_graph=False):
    """Per-structure energies [S], forces [N, 3] and virials [S, 3, 3]
    (W = sum_e r_e (x) dE/dr_e; stress_voigt = -W[voigt]/V) through the
    adjoint path; one launch per kernel for the whole batch.

    The network is differentiated with respect to the image separations r_e
    only; forces and virials then come from dE/dr_e: the gather kernel
    (dE/dR_i = sum_{e of i} dE/dr_rev(e) - dE/dr_e, and the per-centre
    r_e (x) dE/dr_e), summed per structure.  With ``create_graph`` the same
    quantities are built from differentiable ops (force training)."""
    b = desc.batch
    nl = b.nl
    k = nl.nimg
    r = nl.img_r[: 3 * k].view(k, 3).detach().requires_grad_(True)
    x = desc.descriptors(r)
    e_atom = model.atom_energies(x, species)
    e_struct = b.per_structure(e_atom)
    (dedr,) = torch.autograd.grad(e_struct.sum(), [r], create_graph=create_graph)
    out = {"energy": e_struct, "x": x, "atom_energies": e_atom.detach()}
    if create_graph:
        out["forces"] = -adjoint._Gather.apply(dedr, nl)
        if virial:
            w = (r.detach().unsqueeze(2) * dedr.unsqueeze(1)).reshape(k, 9)
            out["virial"] = b.per_structure_images(w).view(b.nstruct, 3, 3)
        return out
    if virial:
        g, vir = dev_so4.gather(nl, dedr, virial=True, img_r=r.detach())
        out["virial"] = b.per_structure(vir.reshape(-1, 9)).view(b.nstruct, 3, 3)
    else:
        g = dev_so4.gather(nl, dedr, img_r=r.detach())
    out["forces"] = -g
    return out


VOIGT = ([0, 1, 2, 0, 0, 1], [0, 1, 2, 1, 2, 2])


def stress_voigt(virial, batch):
    """-W/V in the reference's Voigt order [xx, yy, zz, xy, xz, yz] (eV/A^3),
    the sign of neuralnetwork.py:494-497 applied to rdxdr = sum r (x) dx/dr / V."""
    return -virial[:, VOIGT[0], VOIGT[1]] * batch.inv_vol_struct[:, None]
