"""Device SO(4)/SNAP bispectrum operations on a ``DeviceNeighborList``.

Each function enqueues hand-written gfx950 kernels from libpxff.so on the
current torch stream; tensors are float64 on the GPU.  The math follows the
reference's ``get_bispectrum_components`` (``pyxtal_ff/descriptors/SNAP.py:
1289-1762``; ``SO4.py:1256-1730`` is the same with rfac0 = 1 and Z weights).
"""
import ctypes
from functools import lru_cache

import torch

from . import _native as nat

CUTOFF_IDS = {"cosine": 1, "tanh": 2, "poly1": 3, "poly2": 4, "poly3": 5, "poly4": 6, "exp": 7}


@lru_cache(maxsize=None)
def sizes(twol):
    """dict(nu, nhalf, nz, nb) for 2J = twol."""
    v = [ctypes.c_int() for _ in range(4)]
    nat.check(nat.load().pxff_so4_sizes(int(twol), *[ctypes.byref(x) for x in v]), "pxff_so4_sizes")
    return dict(nu=v[0].value, nhalf=v[1].value, nz=v[2].value, nb=v[3].value)


class So4Config:
    def __init__(self, twol, rcut, rfac0=1.0, cutoff="cosine", normalize_u=False):
        if isinstance(cutoff, str):
            cutoff = CUTOFF_IDS[cutoff]
        self.twol = int(twol)
        self.rcut = float(rcut)
        self.rfac0 = float(rfac0)
        self.cutoff_id = int(cutoff)
        self.normalize_u = bool(normalize_u)
        self._p = nat.So4Params(self.twol, self.rcut, self.rfac0, self.cutoff_id, int(self.normalize_u))
        s = sizes(self.twol)
        self.nu, self.nhalf, self.nz, self.nb = s["nu"], s["nhalf"], s["nz"], s["nb"]

    @property
    def params(self):
        return ctypes.byref(self._p)


def _f64(n, dev):
    return torch.empty(max(int(n), 1), dtype=torch.float64, device=dev)


def ui(nl, cfg, atom_w, site_w, work_order=None, img_r=None, staged=True):
    """U_tot half rows, tile-major: [ceil(N/64), nhalf, 64] complex (float64
    pairs), padding lanes zero.  staged: the per-image values of the sweeps
    computed by a kernel of their own first (pxff_so4_ui_staged, a workspace
    of 6 doubles per image of the list's capacity)."""
    p = nat.ptr
    ntiles = (nl.natoms + 63) // 64
    # every element, padding lanes of the last tile included, is written
    out = torch.empty(max(ntiles * 64 * cfg.nhalf * 2, 1), dtype=torch.float64, device=nl.pos.device)
    r = nl.img_r if img_r is None else img_r.contiguous()
    if staged:
        cap = nl.img_capacity
        stage = torch.empty(max(cap * 6, 1), dtype=torch.float64, device=nl.pos.device)
        nat.check(nat.lib().pxff_so4_ui_staged(cfg.params, nl.natoms, p(nl.img_start), p(nl.img_j), p(r),
                                               p(atom_w), p(site_w), p(work_order), cap, p(stage), p(out),
                                               nat.stream()), "pxff_so4_ui_staged")
        return out
    nat.check(nat.lib().pxff_so4_ui(cfg.params, nl.natoms, p(nl.img_start), p(nl.img_j), p(r), p(atom_w),
                                    p(site_w), p(work_order), p(out), nat.stream()), "pxff_so4_ui")
    return out


@lru_cache(maxsize=None)
def _zlen_per_atom(twol):
    n = ctypes.c_int64(0)
    nat.check(nat.lib().pxff_so4_zlen(int(twol), 64, ctypes.byref(n)), "pxff_so4_zlen")
    return n.value // 64


def zlen(cfg, natoms):
    """Length (float64) of the Z buffer of pxff_so4_zb (pxff_so4_zlen)."""
    return (int(natoms) + 63) // 64 * 64 * _zlen_per_atom(cfg.twol)


def z(nl, cfg, utot):
    """Z from the tile-major U_tot (one kernel launch), tile-major too:
    [ceil(N/64), nz, 64] complex (flat float64 pairs), padding lanes zero;
    the buffer has room for the B row partials after Z."""
    return zb(nl, cfg, utot, want_b=False)[0]


def zb(nl, cfg, utot, want_b=True):
    """Z (tile-major, see z) and B [N, nb]: one Z launch that also writes
    per-row B partials, and a short launch summing them."""
    p = nat.ptr
    zz = _f64(zlen(cfg, nl.natoms), nl.pos.device)
    b = _f64(nl.natoms * cfg.nb, nl.pos.device) if want_b else None
    nat.check(nat.lib().pxff_so4_zb(cfg.params, nl.natoms, p(utot), p(zz), p(b) if want_b else None, nat.stream()),
              "pxff_so4_zb")
    return zz, (b[: nl.natoms * cfg.nb].view(nl.natoms, cfg.nb) if want_b else None)


def bi(nl, cfg, utot, z):
    """B [N, nb] from U_tot and Z."""
    p = nat.ptr
    b = _f64(nl.natoms * cfg.nb, nl.pos.device)
    nat.check(nat.lib().pxff_so4_b(cfg.params, nl.natoms, p(utot), p(z), p(b), nat.stream()), "pxff_so4_b")
    return b[: nl.natoms * cfg.nb].view(nl.natoms, cfg.nb)


def forward(nl, cfg, atom_w, site_w, work_order=None, img_r=None):
    """(B [N, nb], Z, U_tot)."""
    u = ui(nl, cfg, atom_w, site_w, work_order, img_r)
    z, b = zb(nl, cfg, u)
    return b, z, u


def yi(nl, cfg, z, beta):
    p = nat.ptr
    y = _f64(nl.natoms * cfg.nhalf * 2, nl.pos.device)
    beta = beta.contiguous()
    nat.check(nat.lib().pxff_so4_yi(cfg.params, nl.natoms, p(z), p(beta), p(y), nat.stream()), "pxff_so4_yi")
    return y


def deidrj(nl, cfg, atom_w, y, img_r=None, counted=False):
    """dE/dr per image [nimg, 3].  counted: over the list's image capacity
    with the image count read on the device (no host wait; returns
    [capacity, 3], rows at and beyond the count unwritten)."""
    p = nat.ptr
    r = nl.img_r if img_r is None else img_r.contiguous()
    if counted:
        cap = nl.img_capacity
        out = _f64(cap * 3, nl.pos.device)
        count = nl.img_start[nl.natoms:]
        nat.check(nat.lib().pxff_so4_deidrj_counted(cfg.params, cap, p(count), p(nl.img_atom), p(nl.img_j), p(r),
                                                    p(atom_w), p(y), p(out), nat.stream()), "pxff_so4_deidrj_counted")
        return out[: cap * 3].view(cap, 3)
    out = _f64(nl.nimg * 3, nl.pos.device)
    nat.check(nat.lib().pxff_so4_deidrj(cfg.params, nl.nimg, p(nl.img_atom), p(nl.img_j), p(r), p(atom_w), p(y),
                                        p(out), nat.stream()), "pxff_so4_deidrj")
    return out[: nl.nimg * 3].view(nl.nimg, 3)


def gather(nl, dedr, virial=False, img_r=None, forces=False):
    """dE/dR per atom [N, 3] (with ``forces``: the forces -dE/dR, summed
    with the opposite sign in the kernel) and (optionally) the per-centre
    virial [N, 3, 3]."""
    p = nat.ptr
    dev = nl.pos.device
    grad = _f64(nl.natoms * 3, dev)
    vir = _f64(nl.natoms * 9, dev) if virial else None
    r = nl.img_r if img_r is None else img_r.contiguous()
    d = dedr.contiguous()
    fn = "pxff_so4_forces" if forces else "pxff_so4_gather"
    nat.check(getattr(nat.lib(), fn)(nl.natoms, p(nl.img_start), p(nl.img_rev), p(r), p(d), p(grad), p(vir),
                                     nat.stream()), fn)
    g = grad[: nl.natoms * 3].view(nl.natoms, 3)
    return (g, vir[: nl.natoms * 9].view(nl.natoms, 3, 3)) if virial else g


def forces_energy(nl, dedr, e_atom, virial=False):
    """The forces -dE/dR [N, 3] (as ``gather(..., forces=True)``), the
    total energy [1] = sum of e_atom (as the one-segment ``pxff_segment_sum``)
    and (optionally) the per-centre virial [N, 3, 3], in one launch."""
    p = nat.ptr
    dev = nl.pos.device
    f = _f64(nl.natoms * 3, dev)
    vir = _f64(nl.natoms * 9, dev) if virial else None
    e = e_atom.contiguous()
    energy = _f64(1, dev)
    nat.check(nat.lib().pxff_so4_forces_energy(nl.natoms, p(nl.img_start), p(nl.img_rev), p(nl.img_r),
                                               p(dedr.contiguous()), p(f), p(vir), p(e), e.numel(), p(energy),
                                               nat.stream()), "pxff_so4_forces_energy")
    fv = f[: nl.natoms * 3].view(nl.natoms, 3)
    return fv, energy, (vir[: nl.natoms * 9].view(nl.natoms, 3, 3) if virial else None)


def tangent(nl, cfg, atom_w, z, dr, img_r=None):
    """dB [N, nb] = sum_e dB/dr_e . dr_e (forward-mode derivative of B)."""
    p = nat.ptr
    dev = nl.pos.device
    du = _f64((nl.natoms + 63) // 64 * 64 * cfg.nhalf * 2, dev)  # tile-major dU_tot
    db = _f64(nl.natoms * cfg.nb, dev)
    r = nl.img_r if img_r is None else img_r.contiguous()
    d = dr.contiguous()
    nat.check(nat.lib().pxff_so4_tangent(cfg.params, nl.natoms, p(nl.img_start), p(nl.img_j), p(r), p(atom_w),
                                         p(d), p(z), p(du), p(db), nat.stream()), "pxff_so4_tangent")
    return db[: nl.natoms * cfg.nb].view(nl.natoms, cfg.nb)


def tangent_noalloc(nl, cfg, atom_w, z, dr, db, img_r=None):
    """``tangent`` into a caller-owned db [N, nb] with no device allocation
    (pxff_so4_tangent_noalloc: the dU_tot workspace is the library's static
    32 MiB array, so ceil(N/64)*64*nhalf*16 bytes must fit; the 2J tables
    must already exist on the device, e.g. after one ``tangent`` call).
    Calls that overlap on one GPU share the workspace and race."""
    p = nat.ptr
    r = nl.img_r if img_r is None else img_r.contiguous()
    d = dr.contiguous()
    if not (db.is_contiguous() and db.dtype == torch.float64 and db.numel() >= nl.natoms * cfg.nb):
        raise nat.PxffError("db must be a contiguous float64 device tensor of natoms x nb elements")
    nat.check(nat.lib().pxff_so4_tangent_noalloc(cfg.params, nl.natoms, p(nl.img_start), p(nl.img_j), p(r),
                                                 p(atom_w), p(d), p(z), p(db), nat.stream()),
              "pxff_so4_tangent_noalloc")
    return db


def z_atom_major(nl, cfg, z):
    """Atom-major copy [N, nz, 2] of the tile-major Z."""
    za = _f64(nl.natoms * cfg.nz * 2, nl.pos.device)
    nat.check(nat.lib().pxff_so4_z_atom_major(cfg.params, nl.natoms, nat.ptr(z), nat.ptr(za), nat.stream()),
              "pxff_so4_z_atom_major")
    return za


def dbdr(nl, cfg, atom_w, z, img_r=None, atom_major=False):
    """Materialised dB/dr per image [nimg, nb, 3], one workgroup per centre
    (pxff_so4_dbdr_centres), reading the tile-major Z itself or an atom-major
    copy of it (atom_major: fewer cache lines, the same kernel time at 2J=8
    on the 4096-atom benchmark, plus the copy)."""
    p = nat.ptr
    out = _f64(nl.nimg * cfg.nb * 3, nl.pos.device)
    r = nl.img_r if img_r is None else img_r.contiguous()
    zz = z_atom_major(nl, cfg, z) if atom_major else z
    nat.check(nat.lib().pxff_so4_dbdr_centres(cfg.params, nl.natoms, p(nl.img_start), p(nl.img_j), p(r), p(atom_w),
                                              p(zz), 1 if atom_major else 0, p(out), nat.stream()),
              "pxff_so4_dbdr_centres")
    return out


def dbdr_per_image(nl, cfg, atom_w, z, img_r=None):
    """The same dB/dr from the per-image kernel over an atom-major copy of Z
    (pxff_so4_z_atom_major + pxff_so4_dbdr)."""
    p = nat.ptr
    out = _f64(nl.nimg * cfg.nb * 3, nl.pos.device)
    r = nl.img_r if img_r is None else img_r.contiguous()
    za = z_atom_major(nl, cfg, z)
    nat.check(nat.lib().pxff_so4_dbdr(cfg.params, nl.nimg, p(nl.img_atom), p(nl.img_j), p(r), p(atom_w), p(za),
                                      p(out), nat.stream()), "pxff_so4_dbdr")
    return out


def rows(nl, nb, dbdr_img, stress=True):
    """Reference-layout dxdr [nrows, nb, 3] and per-row rdxdr [nrows, nb, 3, 3]
    (= -bstress / V, SNAP.py:219-238)."""
    p = nat.ptr
    dev = nl.pos.device
    dx = _f64(nl.nrows * nb * 3, dev)
    rd = _f64(nl.nrows * nb * 9, dev) if stress else None
    nat.check(nat.lib().pxff_so4_rows(nl.natoms, nb, p(nl.row_start), p(nl.row_img), p(nl.seq), p(nl.nsite),
                                      p(nl.pos), p(nl.img_r), p(dbdr_img), p(nl.inv_vol), p(dx), p(rd),
                                      nat.stream()), "pxff_so4_rows")
    dx = dx[: nl.nrows * nb * 3].view(nl.nrows, nb, 3)
    if stress:
        rd = rd[: nl.nrows * nb * 9].view(nl.nrows, nb, 3, 3)
    return dx, rd


def dxdr_contract(nl, nb, beta, dxdr):
    """dE/dR [N, 3] = sum over rows (i, m) of beta_i . dxdr[(i, m)]."""
    p = nat.ptr
    grad = _f64(nl.natoms * 3, nl.pos.device)
    beta = beta.contiguous()
    nat.check(nat.lib().pxff_dxdr_contract(nl.natoms, nb, p(nl.row_start), p(nl.seq), p(nl.row_rev), p(beta),
                                           p(dxdr), p(grad), nat.stream()), "pxff_dxdr_contract")
    return grad[: nl.natoms * 3].view(nl.natoms, 3)


def dxdr_contract_t(nl, nb, g, dxdr):
    """grad_beta [N, nb] = sum over centre i's rows (i, m) of dxdr[(i, m)] . g[m]
    (the transpose of dxdr_contract)."""
    p = nat.ptr
    out = _f64(nl.natoms * nb, nl.pos.device)
    g = g.contiguous()
    nat.check(nat.lib().pxff_dxdr_contract_t(nl.natoms, nb, p(nl.row_start), p(nl.seq), p(g), p(dxdr), p(out),
                                             nat.stream()), "pxff_dxdr_contract_t")
    return out[: nl.natoms * nb].view(nl.natoms, nb)


def rdxdr_reduce(nl, nb, rdxdr):
    """Per-atom Voigt rdxdr [N, nb, 6] (utilities/__init__.py:226-235)."""
    p = nat.ptr
    out = _f64(nl.natoms * nb * 6, nl.pos.device)
    nat.check(nat.lib().pxff_rdxdr_reduce(nl.natoms, nb, p(nl.row_start), p(rdxdr), p(out), nat.stream()),
              "pxff_rdxdr_reduce")
    return out[: nl.natoms * nb * 6].view(nl.natoms, nb, 6)
