"""Shared engine of the two SO(4) bispectrum classes.

``pyxtal_ff_amd.descriptors.SNAP.SO4_Bispectrum`` (weights dict, self weight
1, rfac0; reference ``pyxtal_ff/descriptors/SNAP.py:10-334``) and
``pyxtal_ff_amd.descriptors.SO4.SO4_Bispectrum`` (weights Z_j, self weight
Z_i, rfac0 = 1; reference ``SO4.py:10-304``) differ only in the weights and
psi; both run the same device path.
"""
import numpy as np
import torch

from .. import so4 as dev_so4
from ..neighbors import DeviceNeighborList
from ..structure import as_arrays

CUTOFF_NAMES = ("cosine", "tanh", "poly1", "poly2", "poly3", "poly4", "exp")


def prefix_count(atom_ids, natoms):
    """The reference indexes its centre arrays by atom id, so only a prefix
    range(k) of atoms works there (SNAP.py:253-261); accept exactly that."""
    if atom_ids is None:
        return natoms
    ids = list(atom_ids)
    if ids != list(range(len(ids))) or len(ids) > natoms:
        raise ValueError("atom_ids must be None or range(k) with k <= len(atoms)")
    return len(ids)


def run_so4(atoms, twol, rcut, rfac0, cutoff_id, normalize_u, derivative, stress, atom_w, site_w,
            atom_ids=None, as_numpy=True):
    """Device SO(4) bispectrum of one structure in the reference's output
    layout (SNAP.py:200-238)."""
    if twol % 2:
        raise NotImplementedError("2J must be even: for odd 2J the reference sizes its output with "
                                  "lmax = 2J//2 and writes past it (SNAP.py:319-334, 1343-1356)")
    pos, cell, numbers, symbols, pbc = as_arrays(atoms)
    n = len(pos)
    k = prefix_count(atom_ids, n)
    vol = abs(float(np.linalg.det(cell)))
    # the image count is settled before the first kernel (a build that
    # overflowed its capacity is redone first, never computed on)
    nl = DeviceNeighborList([pos], [cell], [pbc], rcut).resolve()
    cfg = dev_so4.So4Config(twol, rcut, rfac0, cutoff_id, normalize_u)
    d = nl.pos.device
    aw = torch.as_tensor(np.asarray(atom_w, dtype=np.float64), device=d)
    sw = torch.as_tensor(np.asarray(site_w, dtype=np.float64), device=d)
    b, z, _ = dev_so4.forward(nl, cfg, aw, sw)
    out = {"x": b[:k]}
    if derivative:
        dbi = dev_so4.dbdr(nl, cfg, aw, z)
        dx, rd = dev_so4.rows(nl, cfg.nb, dbi, stress=stress)
        nr = int(nl.row_start[k].item())
        out["dxdr"] = dx[:nr]
        out["seq"] = nl.seq_array()[:nr]
        out["rdxdr"] = rd[:nr] if stress else None
    else:
        out["dxdr"] = None
        if stress:
            # the reference returns -bstress/V of zeros here (bstress is only
            # filled on the derivative branch, SNAP.py:1671-1728)
            out["rdxdr"] = torch.zeros((nl.nrows, cfg.nb, 3, 3), dtype=torch.float64, device=d)
        else:
            out["rdxdr"] = None
    out["elements"] = list(symbols)[:n]
    if as_numpy:
        for key in ("x", "dxdr", "rdxdr", "seq"):
            if key in out and isinstance(out[key], torch.Tensor):
                out[key] = out[key].cpu().numpy()
    return out
