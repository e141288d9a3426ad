"""Shared helpers of the test-suite (comparison metrics, structure access)."""
import numpy as np


def rel_err(a, b):
    """Norm-wise relative error ||a - b|| / max(||b||, tiny) -- the bar for
    arrays that contain near-cancelling entries (the (i, i) rows)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / max(nb, 1e-300))


def max_rel_elem(a, b, floor):
    """max |a - b| / max(|b|, floor) elementwise."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


def struct_arrays(s):
    return s.positions, s.cell, s.pbc, s.numbers, s.symbols


def contract_rows(d, beta):
    """Forces -sum over the seq rows (i, m) of beta_i . dxdr[(i, m)] of a
    reference-layout descriptor dict -- the reference's contraction
    einsum("ik,ijkl->jl") over the dense dxdr (neuralnetwork.py:484-492),
    taken row by row."""
    seq = np.asarray(d["seq"])
    n = d["x"].shape[0]
    f = np.zeros((n, 3))
    np.add.at(f, seq[:, 1], -np.einsum("rb,rbk->rk", np.asarray(beta)[seq[:, 0]], d["dxdr"]))
    return f
