"""Device SO(3) power spectrum operations on a ``DeviceNeighborList``.

Host side: the radial basis tables the reference recomputes on every call
(``pyxtal_ff/descriptors/SO3.py:466-502``: W = S^-1/2 by inverse +
eigendecomposition, g_n on the Gauss-Chebyshev nodes, the node weights and
the r^2 exp(-alpha r^2) sqrt(1 - x^2) factors) are built once per parameter
set in NumPy and uploaded.  Device side: ``pxff_so3_*`` kernels (expansion,
power spectrum, plain dP/dr, adjoint dE/dr, tangent).
"""
import ctypes
from functools import lru_cache

import numpy as np
import torch

from . import _native as nat

CUTOFF_IDS = {"cosine": 1, "tanh": 2, "poly1": 3, "poly2": 4, "poly3": 5, "poly4": 6, "exp": 7, "unity": 8}
RWORK = 256 * 64 + 1  # |r| bins x chunks + count: the fixed part of pxff_so3_ctot_pairs' workspace


def W(nmax):
    """S^-1/2 of the overlap of the (rc - r)^(a+2) functions (SO3.py:466-479)."""
    arr = np.zeros((nmax, nmax), np.float64)
    for a in range(1, nmax + 1):
        t1 = (2 * a + 5) * (2 * a + 6) * (2 * a + 7)
        for b in range(1, a + 1):
            t2 = (2 * b + 5) * (2 * b + 6) * (2 * b + 7)
            arr[a - 1, b - 1] = np.sqrt(t1 * t2) / (5 + a + b) / (6 + a + b) / (7 + a + b)
            arr[b - 1, a - 1] = arr[a - 1, b - 1]
    sinv = np.linalg.inv(arr)
    ev, V = np.linalg.eig(sinv)
    return np.dot(np.dot(V, np.diag(np.sqrt(ev))), np.linalg.inv(V))


@lru_cache(maxsize=32)
def radial_tables(nmax, lmax, rcut, alpha):
    """(G [nmax, NQ], r_q [NQ], norm_l [lmax+1]) -- SO3.py:209-210, 495-502, 589-617."""
    nq = (nmax + lmax + 1) * 10
    x = np.cos((2 * np.arange(1, nq + 1) - 1) * np.pi / 2 / nq)
    weight = np.pi / nq * rcut / 2
    Q = rcut / 2 * (x + 1)
    w = W(nmax)
    G = np.zeros((nmax, nq))
    for n in range(1, nmax + 1):
        s = 0.0
        for a in range(1, nmax + 1):
            phi = (rcut - Q) ** (a + 2) / np.sqrt(2 * rcut ** (2 * a + 7) / (2 * a + 5) / (2 * a + 6) / (2 * a + 7))
            s = s + w[n - 1, a - 1] * phi
        G[n - 1] = s * weight
    G *= Q ** 2 * np.exp(-alpha * Q ** 2) * np.sqrt(1 - x ** 2)
    ls = np.arange(lmax + 1)
    lnorm = np.sqrt(2 * np.sqrt(2) * np.pi / np.sqrt(2 * ls + 1))
    return G, Q, lnorm


class So3Config:
    def __init__(self, nmax, lmax, rcut, alpha=2.0, cutoff="cosine", weight_on=False, device="cuda"):
        if isinstance(cutoff, str):
            cutoff = CUTOFF_IDS[cutoff]
        self.nmax, self.lmax = int(nmax), int(lmax)
        self.rcut, self.alpha = float(rcut), float(alpha)
        self.cutoff_id = int(cutoff)
        self.weight_on = bool(weight_on)
        self.nq = (self.nmax + self.lmax + 1) * 10
        self.ntri = self.nmax * (self.nmax + 1) // 2
        self.ncoef = self.ntri * (self.lmax + 1)
        self.csz = self.nmax * (self.lmax + 1) * (2 * self.lmax + 1)
        G, Q, ln = radial_tables(self.nmax, self.lmax, self.rcut, self.alpha)
        dev = torch.device(device)
        self.G = torch.as_tensor(np.ascontiguousarray(G), dtype=torch.float64, device=dev)
        self.rq = torch.as_tensor(Q, dtype=torch.float64, device=dev)
        self.lnorm = torch.as_tensor(ln, dtype=torch.float64, device=dev)
        self._p = nat.So3Params(self.nmax, self.lmax, self.rcut, self.alpha, self.cutoff_id, self.nq,
                                int(self.weight_on))
        self._table = None

    @property
    def params(self):
        return ctypes.byref(self._p)

    def radial_table(self):
        """The tabulated radial integrals (pxff_so3_radial_table) for
        nmax, lmax <= 6, built on the device at first use; None otherwise."""
        if self.nmax > 6 or self.lmax > 6:
            return None
        if self._table is None:
            nk, n = ctypes.c_int(), ctypes.c_int64()
            nat.check(nat.lib().pxff_so3_radial_table_size(self.params, ctypes.byref(nk), ctypes.byref(n)),
                      "pxff_so3_radial_table_size")
            tab = torch.empty(n.value, dtype=torch.float64, device=self.G.device)
            scratch = torch.empty(n.value, dtype=torch.float64, device=self.G.device)
            p = nat.ptr
            nat.check(nat.lib().pxff_so3_radial_table(self.params, p(self.G), p(self.rq), p(self.lnorm), p(scratch),
                                                      p(tab), nat.stream()), "pxff_so3_radial_table")
            self._table = tab
        return self._table


def _f64(n, dev):
    return torch.empty(max(int(n), 1), dtype=torch.float64, device=dev)


def ctot(nl, cfg, atom_z, img_r=None, dr=None, radial=False, use_table=True):
    """Per-centre expansion [N, nmax, lmax+1, 2lmax+1] complex (as float64
    pairs); with dr, its tangent along the per-image displacements.  With
    radial=True (plain expansion only) also the per-image radial integrals
    [nimg, lmax+1, 2, nmax] for the backward: returns (ctot, radial).  With
    the list's own separations and nmax, lmax <= 6 the integrals come from
    the configuration's table (use_table) or the quadrature."""
    p = nat.ptr
    out = _f64(nl.natoms * cfg.csz * 2, nl.pos.device)
    r = nl.img_r if img_r is None else img_r.contiguous()
    d = None if dr is None else dr.contiguous()
    # the plain expansion always goes through the radial store (the
    # per-(l, m) accumulation reads it); it is returned only when asked for
    keep = radial
    radial = radial or d is None
    rad = _f64(nl.img_capacity * (cfg.lmax + 1) * 2 * cfg.nmax, nl.pos.device) if radial else None
    if radial and d is None and r.data_ptr() == nl.img_r.data_ptr():
        # the list's own separations: an image and its reverse share |r|,
        # so the radial integrals are evaluated once per pair
        # workspace of the |r| ordering of the pairs (scheduling only)
        cap = nl.img_capacity
        work = torch.empty(RWORK + cap, dtype=torch.int64, device=nl.pos.device)
        tab = cfg.radial_table() if use_table else None
        nat.check(nat.lib().pxff_so3_ctot_pairs(cfg.params, nl.natoms, p(nl.img_start), p(nl.img_j), p(r),
                                                p(nl.img_rev), p(atom_z), p(cfg.G), p(cfg.rq), p(cfg.lnorm),
                                                p(out), p(rad), p(work), cap, p(tab) if tab is not None else None,
                                                nat.stream()), "pxff_so3_ctot_pairs")
        return (out, rad) if keep else out
    nat.check(nat.lib().pxff_so3_ctot(cfg.params, nl.natoms, p(nl.img_start), p(nl.img_j), p(r), p(atom_z),
                                      p(cfg.G), p(cfg.rq), p(cfg.lnorm), p(d), p(out), p(rad), nat.stream()),
              "pxff_so3_ctot")
    return (out, rad) if keep else out


def power(nl, cfg, ct, dct=None):
    p = nat.ptr
    x = _f64(nl.natoms * cfg.ncoef, nl.pos.device)
    nat.check(nat.lib().pxff_so3_power(cfg.params, nl.natoms, p(ct), p(dct), p(x), nat.stream()), "pxff_so3_power")
    return x[: nl.natoms * cfg.ncoef].view(nl.natoms, cfg.ncoef)


def forward(nl, cfg, atom_z, img_r=None, radial=False):
    """(x, ctot) or, with radial=True, (x, ctot, radial integrals)."""
    if radial:
        ct, rad = ctot(nl, cfg, atom_z, img_r, radial=True)
        return power(nl, cfg, ct), ct, rad
    ct = ctot(nl, cfg, atom_z, img_r)
    return power(nl, cfg, ct), ct


def adjoint_a(nl, cfg, ct, beta):
    p = nat.ptr
    a = _f64(nl.natoms * cfg.csz * 2, nl.pos.device)
    beta = beta.contiguous()
    nat.check(nat.lib().pxff_so3_adjoint_a(cfg.params, nl.natoms, p(ct), p(beta), p(a), nat.stream()),
              "pxff_so3_adjoint_a")
    return a


def deidrj(nl, cfg, atom_z, a, img_r=None, radial=None, counted=False):
    """dE/dr per image [nimg, 3]; radial: the forward's radial integrals (or
    None to recompute them).  counted: over the list's image capacity with
    the image count read on the device (no host wait; returns [capacity,
    3], rows at and beyond the count unwritten)."""
    p = nat.ptr
    r = nl.img_r if img_r is None else img_r.contiguous()
    if counted:
        cap = nl.img_capacity
        out = _f64(cap * 3, nl.pos.device)
        count = nl.img_start[nl.natoms:]
        nat.check(nat.lib().pxff_so3_deidrj_counted(cfg.params, cap, p(count), p(nl.img_atom), p(nl.img_j), p(r),
                                                    p(atom_z), p(cfg.G), p(cfg.rq), p(cfg.lnorm), p(a), p(radial),
                                                    p(out), nat.stream()), "pxff_so3_deidrj_counted")
        return out[: cap * 3].view(cap, 3)
    out = _f64(nl.nimg * 3, nl.pos.device)
    nat.check(nat.lib().pxff_so3_deidrj(cfg.params, nl.nimg, p(nl.img_atom), p(nl.img_j), p(r), p(atom_z), p(cfg.G),
                                        p(cfg.rq), p(cfg.lnorm), p(a), p(radial), p(out), nat.stream()),
              "pxff_so3_deidrj")
    return out[: nl.nimg * 3].view(nl.nimg, 3)


def dpdr(nl, cfg, atom_z, ct, img_r=None, radial=None):
    """Materialised dP/dr per image [nimg, ncoef, 3]."""
    p = nat.ptr
    out = _f64(nl.nimg * cfg.ncoef * 3, nl.pos.device)
    r = nl.img_r if img_r is None else img_r.contiguous()
    nat.check(nat.lib().pxff_so3_dpdr(cfg.params, nl.nimg, p(nl.img_atom), p(nl.img_j), p(r), p(atom_z), p(cfg.G),
                                      p(cfg.rq), p(cfg.lnorm), p(ct), p(radial), p(out), nat.stream()),
              "pxff_so3_dpdr")
    return out


class SO3State:
    """Neighbour list, config and atomic numbers; store_radial keeps the
    per-image radial integrals (2 nmax (lmax+1) doubles per image) from the
    forward for the backward instead of recomputing the quadrature."""

    def __init__(self, nl, cfg, atom_z, store_radial=True):
        self.nl, self.cfg, self.atom_z = nl, cfg, atom_z
        self.store_radial = store_radial


class _PFromR(torch.autograd.Function):
    """x [N, ncoef] as a function of the separations r [nimg, 3]; backward =
    adjoint (A = S(dE/dx) ctot, one sweep of the angular part per image),
    double backward = tangent of x.  Per-image expansions are recomputed,
    never stored; the radial integrals are kept when store_radial."""

    @staticmethod
    def forward(ctx, r, state):
        if state.store_radial:
            x, ct, rad = forward(state.nl, state.cfg, state.atom_z, img_r=r, radial=True)
        else:
            (x, ct), rad = forward(state.nl, state.cfg, state.atom_z, img_r=r), None
        ctx.state, ctx.ct, ctx.rad = state, ct, rad
        ctx.save_for_backward(r)
        return x

    @staticmethod
    def backward(ctx, beta):
        (r,) = ctx.saved_tensors
        return _SO3DEdR.apply(beta.contiguous(), r, ctx.ct, ctx.state, ctx.rad), None


class _SO3DEdR(torch.autograd.Function):
    @staticmethod
    def forward(ctx, beta, r, ct, state, rad=None):
        a = adjoint_a(state.nl, state.cfg, ct, beta)
        ctx.state, ctx.ct = state, ct
        ctx.save_for_backward(r)
        return deidrj(state.nl, state.cfg, state.atom_z, a, img_r=r, radial=rad)

    @staticmethod
    def backward(ctx, g):
        (r,) = ctx.saved_tensors
        st = ctx.state
        dct = ctot(st.nl, st.cfg, st.atom_z, img_r=r, dr=g.contiguous())
        return power(st.nl, st.cfg, ctx.ct, dct), None, None, None, None


def power_spectrum(r, state):
    return _PFromR.apply(r, state)
