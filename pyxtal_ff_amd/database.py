"""Descriptor records of a training set, batched on the GPU.

The reference's ``Database`` (``pyxtal_ff/utilities/__init__.py:14-254``)
computes one record per structure -- in a ``multiprocessing.Pool`` --
and stores it in a ``shelve`` file under the key ``str(index)``:

    x [N, nc], dxdr [P, nc, 3], seq [P, 2] (indices local to the structure),
    rdxdr [N, nc, 6] (per-atom sum of the rows, Voigt xx yy zz xy xz yz),
    elements, energy, force, stress, group

(``Database.compute``, ``:166-254``).  Here a chunk of structures is one
neighbour list and every kernel (descriptor, plain dB/dr, the reference's
row layout, the per-atom rdxdr reduction) runs once per chunk; the records
are then cut out of the chunk's arrays.  Same keys, shapes and dtypes; no
base potential (ZBL) is subtracted.
"""
import shelve

import numpy as np
import torch

from . import batch as bt
from . import so3 as dso3
from . import so4 as dso4
from .calculator import make_descriptor


def _plain_arrays(desc):
    """x, dxdr rows and per-atom Voigt rdxdr of a batch (device arrays)."""
    nl = desc.batch.nl.resolve()  # count settled before any kernel
    if isinstance(desc, bt.BatchedSO3):
        x, ct, rad = dso3.forward(nl, desc.cfg, desc.state.atom_z, radial=True)
        nf = desc.cfg.ncoef
        dx, rd = dso4.rows(nl, nf, dso3.dpdr(nl, desc.cfg, desc.state.atom_z, ct, radial=rad), stress=True)
    else:
        x, z, _ = dso4.forward(nl, desc.cfg, desc.atom_w, desc.site_w)
        nf = desc.cfg.nb
        dx, rd = dso4.rows(nl, nf, dso4.dbdr(nl, desc.cfg, desc.atom_w, z), stress=True)
    rv = dso4.rdxdr_reduce(nl, nf, rd)
    return x, dx, rv


def compute_records(function, data, chunk_atoms=20000, device="cuda"):
    """Records (dicts in the reference's format) of a list of entries
    {'structure', 'energy', 'force', 'stress', 'group'}.  function: the
    reference's descriptor dict ('type', 'Rc', 'parameters', 'weights',
    'cutoff', and the 'force' / 'stress' switches: without force there is
    no dxdr and no seq, without stress no rdxdr).  Structures are processed
    in chunks of about chunk_atoms atoms."""
    out = []
    k = 0
    while k < len(data):
        n = 0
        j = k
        while j < len(data) and (j == k or n + len(data[j]["structure"]) <= chunk_atoms):
            n += len(data[j]["structure"])
            j += 1
        structs = [d["structure"] for d in data[k:j]]
        sb = bt.StructureBatch(structs, float(function["Rc"]), device=device)
        desc = make_descriptor(sb, function)
        force = function.get("force", True)
        stress = force and function.get("stress", True)
        nl = sb.nl
        a0 = sb.builder.atom_start_host
        if force:
            x, dx, rv = _plain_arrays(desc)
            dx = dx.cpu().numpy()
            rv = rv.cpu().numpy()
            seq = nl.seq_array().cpu().numpy()
            row_start = nl.row_start.cpu().numpy()
        else:
            x = desc.descriptors(nl.img_r[: 3 * nl.nimg].view(-1, 3))
        x = x.detach().cpu().numpy()
        for s, entry in enumerate(data[k:j]):
            lo, hi = int(a0[s]), int(a0[s + 1])
            rec = {"x": x[lo:hi].copy(), "dxdr": None, "rdxdr": None, "elements": list(sb.symbols[lo:hi])}
            if force:
                r0, r1 = int(row_start[lo]), int(row_start[hi])
                rec["dxdr"] = dx[r0:r1].copy()
                rec["seq"] = seq[r0:r1] - lo
                if stress:
                    rec["rdxdr"] = rv[lo:hi].copy()
            rec["energy"] = np.asarray(entry["energy"])
            rec["force"] = np.asarray(entry["force"])
            rec["stress"] = None if entry.get("stress") is None else np.asarray(entry["stress"])
            rec["group"] = entry.get("group")
            out.append(rec)
        k = j
        del sb, desc, nl
        torch.cuda.empty_cache()
    return out


class Database:
    """shelve-backed record store with the reference's key scheme
    (``str(index)``) and ``add`` computing the records in batched launches."""

    def __init__(self, name):
        self.name = name
        self.database = shelve.open(self.name)
        self.length = len(self.database)

    def __len__(self):
        return len(self.database)

    def __getitem__(self, index):
        if isinstance(index, int) and 0 <= index < len(self):
            return self.database[str(index)]
        raise IndexError("index out of range")

    def append(self, value):
        self.database[str(len(self))] = value
        self.length = len(self)

    def add(self, function, data, chunk_atoms=20000):
        """Compute and store the records of the entries of data
        (``Database.add``, :121-164, one GPU batch per chunk instead of a
        process pool)."""
        n = function.get("N")
        items = data if n is None else data[:n]
        for rec in compute_records(function, items, chunk_atoms):
            self.append(rec)

    def close(self):
        self.database.close()
