"""Adjoint energy / force / virial path behind ``torch.autograd.Function``.

B (the bispectrum, [N, nb]) is exposed as a differentiable function of the
per-image separations r_e.  Its backward contracts dE/dB into the per-atom Y
array (``pxff_so4_yi``) and sweeps dU/dr once per image
(``pxff_so4_deidrj``), so dB/dR is never materialised; the force/virial
contraction this replaces is ``pyxtal_ff/models/neuralnetwork.py:484-497,
657-677`` over the reference's dxdr / rdxdr.  The backward is itself a
Function whose own backward is the forward-mode derivative of B
(``pxff_so4_tangent``), which is what force-matched training needs (the
gradient of a force loss with respect to the network parameters).

Second derivatives with respect to positions (Hessian terms) are not
provided: the double backward returns no gradient for r.  Per-image U / dU
arrays are never stored; with ``store_z=False`` Z is recomputed from U_tot in
the backward instead of being kept between forward and backward.
"""
import torch

from . import so4 as dev_so4


class SO4State:
    """Everything the kernels need besides r: neighbour list, config, weights."""

    def __init__(self, nl, cfg, atom_w, site_w, store_z=True, work_order=None):
        self.nl = nl
        self.cfg = cfg
        self.atom_w = atom_w
        self.site_w = site_w
        self.store_z = store_z
        self.work_order = work_order


class _BFromR(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, state):
        b, z, u = dev_so4.forward(state.nl, state.cfg, state.atom_w, state.site_w, state.work_order, img_r=r)
        ctx.state = state
        if state.store_z:
            ctx.z, ctx.u = z, None
        else:
            ctx.z, ctx.u = None, u
        ctx.save_for_backward(r)
        return b

    @staticmethod
    def backward(ctx, beta):
        (r,) = ctx.saved_tensors
        st = ctx.state
        z = ctx.z
        if z is None:
            z, _ = dev_so4.zb(st.nl, st.cfg, ctx.u, want_b=False)
        dedr = _DEdR.apply(beta.contiguous(), r, z, st)
        return dedr, None


class _DEdR(torch.autograd.Function):
    @staticmethod
    def forward(ctx, beta, r, z, state):
        y = dev_so4.yi(state.nl, state.cfg, z, beta)
        dedr = dev_so4.deidrj(state.nl, state.cfg, state.atom_w, y, img_r=r)
        ctx.state = state
        ctx.z = z
        ctx.save_for_backward(r)
        return dedr

    @staticmethod
    def backward(ctx, g):
        (r,) = ctx.saved_tensors
        st = ctx.state
        db = dev_so4.tangent(st.nl, st.cfg, st.atom_w, ctx.z, g.contiguous(), img_r=r)
        return db, None, None, None


def bispectrum(r, state):
    """B [N, nb] as a differentiable function of the separations r [nimg, 3]."""
    return _BFromR.apply(r, state)


class _Separations(torch.autograd.Function):
    """r_e = pos[j] - pos[i] + (lattice shift of e); backward
    dE/dR_i = sum_{e of i} (dE/dr_rev(e) - dE/dr_e) by the deterministic
    gather kernel (pxff_so4_gather) instead of scatter-adds."""

    @staticmethod
    def forward(ctx, pos, nl):
        ctx.nl = nl
        k = nl.nimg
        if pos.data_ptr() == nl.pos.data_ptr():
            return nl.img_r[: 3 * k].view(k, 3).clone()
        i = nl.img_atom[:k]
        j = nl.img_j[:k]
        return pos.index_select(0, j) - pos.index_select(0, i) + _shift(nl)

    @staticmethod
    def backward(ctx, g):
        return _Gather.apply(g, ctx.nl), None


class _Gather(torch.autograd.Function):
    """dE/dR from dE/dr (pxff_so4_gather); its own backward is the transpose
    dr_e = G[j] - G[i], so forces stay differentiable (force training)."""

    @staticmethod
    def forward(ctx, dedr, nl):
        ctx.nl = nl
        return dev_so4.gather(nl, dedr.contiguous())

    @staticmethod
    def backward(ctx, G):
        nl = ctx.nl
        k = nl.nimg
        return G.index_select(0, nl.img_j[:k]) - G.index_select(0, nl.img_atom[:k]), None


def _shift(nl):
    if getattr(nl, "_shift", None) is None:
        k = nl.nimg
        base = nl.img_r[: 3 * k].view(k, 3)
        nl._shift = base - (nl.pos.index_select(0, nl.img_j[:k]) - nl.pos.index_select(0, nl.img_atom[:k]))
    return nl._shift


def separations(nl, pos, strain=None):
    """r_e = pos[j] - pos[i] + T.cell (+ r_e . strain) as a differentiable
    function of positions (and of a homogeneous strain, whose gradient is
    the virial sum_e r_e (x) dE/dr_e).  With pos equal to the positions the
    list was built from, r_e are the list's (bit-exact) separations."""
    r = _Separations.apply(pos, nl)
    if strain is not None:
        r = r + r @ strain
    return r
