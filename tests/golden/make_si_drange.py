"""Write tests/golden/si_drange.npz from the reference's example artefacts.

Inputs (read as data, nothing executed):
  examples/Si-LAMMPS/Si-snap-zbl/ase.db          SQLite; the 239 Si structures
      (214 training + 25 test) the shipped models were trained on;
      positions/cell/numbers are float64/int32 blobs; the data blob holds
      the DFT energy and forces (the training labels).
  examples/Si-LAMMPS/Si-snap-zbl/NN_weights.txt  Scale0 = per-component min,
      Scale1 = max - min of the SNAP descriptor (2J=6, rc=5.0, w_Si=1.0,
      rfac0=0.99363, cosine) over the training set, printed with 13 decimals
      (pyxtal_ff/models/neuralnetwork.py:755-774).
  examples/Si-LAMMPS/Si-so3/NN_weights.txt       the same for SO3 (nmax 3,
      lmax 4, rc 5.0, alpha 2.0).
Usage: python tests/golden/make_si_drange.py [/path/to/reference]
"""
import json
import os
import sqlite3
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_scales(path):
    lines = open(path).read().split("\n")
    out = {}
    for key in ("Scale0", "Scale1"):
        i = [k for k, l in enumerate(lines) if l.startswith("# " + key)][0]
        vals = []
        for l in lines[i + 1:]:
            if not l.strip():
                break
            vals += [float(v) for v in l.split()]
        out[key] = np.array(vals)
    return out


def main():
    db = os.path.join(REF, "examples/Si-LAMMPS/Si-snap-zbl/ase.db")
    con = sqlite3.connect("file:%s?mode=ro" % db, uri=True)
    rows = con.execute("select id, numbers, positions, cell, pbc, data from systems order by id").fetchall()
    natoms = np.array([len(np.frombuffer(r[1], dtype=np.int32)) for r in rows], dtype=np.int64)
    pos = np.concatenate([np.frombuffer(r[2], dtype=np.float64).reshape(-1, 3) for r in rows])
    cells = np.stack([np.frombuffer(r[3], dtype=np.float64).reshape(3, 3) for r in rows])
    pbc = np.array([[(r[4] >> k) & 1 for k in range(3)] for r in rows], dtype=bool)
    numbers = np.concatenate([np.frombuffer(r[1], dtype=np.int32) for r in rows]).astype(np.int64)
    # DFT labels of each structure: the "data" blob is an 8-byte length then
    # JSON with "energy" (eV) and "force" (eV/A) -- the training targets
    labels = [json.loads(bytes(r[5])[8:].decode()) for r in rows]
    energy = np.array([d["energy"] for d in labels], dtype=np.float64)
    forces = np.concatenate([np.asarray(d["force"], dtype=np.float64).reshape(-1, 3) for d in labels])
    snap = parse_scales(os.path.join(REF, "examples/Si-LAMMPS/Si-snap-zbl/NN_weights.txt"))
    so3 = parse_scales(os.path.join(REF, "examples/Si-LAMMPS/Si-so3/NN_weights.txt"))
    np.savez_compressed(os.path.join(HERE, "si_drange.npz"), natoms=natoms, positions=pos, cells=cells, pbc=pbc,
                        numbers=numbers, ntrain=np.int64(214), snap_min=snap["Scale0"], snap_range=snap["Scale1"],
                        so3_min=so3["Scale0"], so3_range=so3["Scale1"], energy=energy, forces=forces)
    print("wrote", len(rows), "structures,", len(pos), "atoms")


if __name__ == "__main__":
    main()
