"""Copy the reference's shipped LAMMPS ML-IAP export files (data: network
weights and descriptor parameters of the two example Si potentials) into
tests/golden/lammps/.

Inputs (read as data, nothing executed):
  examples/Si-LAMMPS/Si-snap-zbl/{NN_weights,DescriptorParam}.txt  SNAP 2J=6 potential
  examples/Si-LAMMPS/Si-so3/{NN_weights,DescriptorParam}.txt       SO3 nmax 3 lmax 4 potential
The energies printed in examples/Si-LAMMPS/README.md are not usable as
vectors: validate.py as shipped draws a in [4.5, 5.5), while the printed
energies (-221.7 .. -227.2 eV for 64 atoms) and stresses (~16 GPa) belong to
cells near a = 4.5-4.6 with both shipped potentials, so the table came from
another revision of the script.
Usage: python tests/golden/make_lammps_fixtures.py [/path/to/reference]
"""
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lammps")


def main():
    os.makedirs(OUT, exist_ok=True)
    base = os.path.join(REF, "examples", "Si-LAMMPS")
    for model in ("Si-snap-zbl", "Si-so3"):
        for name in ("NN_weights.txt", "DescriptorParam.txt"):
            shutil.copyfile(os.path.join(base, model, name), os.path.join(OUT, model + "_" + name))


if __name__ == "__main__":
    main()
