"""The E+F step captured as one HIP graph (FusedEFGraph: neighbour-list
rebuild plus the fused adjoint kernels) against the eager fused path: atom
energies and forces bit for bit, energy and virial to rounding (their sums
run in another order), over replays with new positions and eager work
between the replays; an image capacity that is too small is detected after
the replay and the step redone with a larger one."""
import numpy as np
import pytest
import torch

from pyxtal_ff_amd import adjoint, potential, workloads
from pyxtal_ff_amd import so4 as dso4
from pyxtal_ff_amd.neighbors import NeighborListBuilder
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _setup(cap=None):
    s = workloads.diamond_si(reps=3, sigma=0.08, seed=4)
    n = len(s)
    cfg = dso4.So4Config(6, 4.9, 0.99363, "cosine")
    pos = torch.as_tensor(s.positions, dtype=torch.float64, device="cuda")
    w = torch.ones(n, dtype=torch.float64, device="cuda")
    model = potential.NNPotential(["Si"], cfg.nb, (16, 16), seed=2).cuda()
    sp = potential.species_atom(s.symbols, ["Si"], "cuda")
    b = NeighborListBuilder([n], [s.cell], [s.pbc], 4.9)
    if cap is not None:
        b.img_capacity = cap
    return s, cfg, pos, w, model, sp, b


def _eager(s, cfg, pos, w, model, sp):
    nl = NeighborListBuilder([len(s)], [s.cell], [s.pbc], 4.9).build(pos)
    return potential.energy_forces_fused(nl, adjoint.SO4State(nl, cfg, w, w), model, sp, virial=True)


def _np(t):
    return t.detach().cpu().numpy()


def test_graph_step_matches_eager_over_replays():
    s, cfg, pos, w, model, sp, b = _setup()
    g = potential.FusedEFGraph(b, cfg, (w, w), model, sp, pos, virial=True)
    rng = np.random.default_rng(0)
    for it in range(3):
        p = pos + 0.02 * torch.as_tensor(rng.standard_normal(tuple(pos.shape)), device="cuda")
        out = g(p)
        got = {k: v.clone() for k, v in out.items()}
        ref = _eager(s, cfg, p, w, model, sp)  # eager work between the replays
        print("replay", it, "max |dF|", float((got["forces"] - ref["forces"]).abs().max()), flush=True)
        assert torch.equal(got["atom_energies"], ref["atom_energies"])
        assert torch.equal(got["forces"], ref["forces"])
        assert rel_err(_np(got["energy"]), _np(ref["energy"])) < 1e-14
        assert rel_err(_np(got["virial"]), _np(ref["virial"])) < 1e-13
    assert g.captures == 1


def test_graph_step_grows_a_small_capacity():
    s, cfg, pos, w, model, sp, b = _setup(cap=100)
    g = potential.FusedEFGraph(b, cfg, (w, w), model, sp, pos)
    out = g(pos)
    assert g.captures == 2 and b.img_capacity > 100
    ref = _eager(s, cfg, pos, w, model, sp)
    assert torch.equal(out["forces"], ref["forces"])
    assert torch.equal(out["atom_energies"], ref["atom_energies"])


def _so3_setup(s, nmax=6, lmax=6, rc=5.0):
    from pyxtal_ff_amd import so3 as dso3
    n = len(s)
    cfg = dso3.So3Config(nmax, lmax, rc, 2.0)
    pos = torch.as_tensor(s.positions, dtype=torch.float64, device="cuda")
    z = torch.as_tensor(s.numbers, dtype=torch.int64, device="cuda")
    el = ["O", "Si"]
    model = potential.NNPotential(el, cfg.ncoef, (16, 16), seed=2).cuda()
    sp = potential.species_atom(s.symbols, el, "cuda")
    return cfg, pos, z, model, sp, NeighborListBuilder([n], [s.cell], [s.pbc], rc)


def test_so3_graph_step_matches_eager_and_autograd():
    """SO(3) 6/6 E+F step (amorphous SiO2, 300 atoms) as one graph replay:
    bit for bit the eager fused step over replays with new positions; forces
    equal the autograd adjoint path's."""
    from pyxtal_ff_amd import batch as bt
    from pyxtal_ff_amd import so3 as dso3
    s = workloads.amorphous_sio2(100, 200, seed=3)
    cfg, pos, z, model, sp, b = _so3_setup(s)
    g = potential.FusedEFGraph(b, cfg, (z,), model, sp, pos, virial=True)
    rng = np.random.default_rng(1)
    for it in range(3):
        p = pos + 0.02 * torch.as_tensor(rng.standard_normal(tuple(pos.shape)), device="cuda")
        got = {k: v.clone() for k, v in g(p).items()}
        nl = NeighborListBuilder([len(s)], [s.cell], [s.pbc], 5.0).build(p)
        ref = potential.energy_forces_fused(nl, dso3.SO3State(nl, cfg, z), model, sp, virial=True)
        print("so3 replay", it, float((got["forces"] - ref["forces"]).abs().max()), flush=True)
        assert torch.equal(got["atom_energies"], ref["atom_energies"])
        assert torch.equal(got["forces"], ref["forces"])
        assert rel_err(_np(got["energy"]), _np(ref["energy"])) < 1e-14
        assert rel_err(_np(got["virial"]), _np(ref["virial"])) < 1e-13
    assert g.captures == 1
    # the autograd adjoint (torch network) on the last positions
    sb = bt.StructureBatch([s], 5.0)
    sb.pos = p
    sb.nl = sb.builder.build(p).resolve()
    d = bt.BatchedSO3(sb, nmax=6, lmax=6, rcut=5.0, alpha=2.0)
    o = bt.energy_forces_virial(d, model, potential.species_index(s.symbols, ["O", "Si"], "cuda"), virial=True)
    assert rel_err(_np(o["forces"]), _np(got["forces"])) < 1e-10
    assert rel_err(_np(o["virial"][0]), _np(got["virial"])) < 1e-10
