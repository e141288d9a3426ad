"""The benchmark workloads themselves against the oracle:

  * the 1024-structure Si/O batch (8-256 atoms, SNAP 2J=8, weights Si 1.0 /
    O 0.8) in one batched launch: x, energies and forces of a 20-structure
    sample (the smallest cells, many images per pair, included) against the
    oracle's descriptors and dxdr contraction;
  * force-matched training on 256 of those structures at full size: the
    captured step replays eager steps, the parameter gradient is a central
    difference of the loss, and the teacher labels equal the oracle's;
  * SO3 nmax = lmax = 6 on amorphous SiO2 (2.2 g/cm^3, 1.6 A minimum
    distance): a 300-atom cell against the oracle for x, dxdr and rdxdr,
    and the 12,000-atom cell (x of 300 centres against the oracle, adjoint
    forces against the plain dP/dr forces, momentum conservation)."""
import numpy as np
import pytest
import torch

from pyxtal_ff_amd import batch as bt
from pyxtal_ff_amd import potential, training, workloads
from pyxtal_ff_amd.descriptors.SO3 import SO3
from tests.helpers import contract_rows, rel_err
from tests.test_gpu_so4 import check_layout

pytestmark = pytest.mark.gpu

W = {"Si": 1.0, "O": 0.8}
EL = ["O", "Si"]


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c3_structs():
    return workloads.random_sio_batch(1024, seed=0)


def _oracle_e_f(oracle_mod, s, model_cpu, lmax=4, rc=4.9):
    d = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, W, lmax=lmax, rcut=rc, derivative=True,
                                  stress=True)
    sp = potential.species_index(s.symbols, EL, "cpu")
    x = torch.as_tensor(d["x"]).requires_grad_(True)
    ea = model_cpu.atom_energies(x, sp)
    (beta,) = torch.autograd.grad(ea.sum(), [x])
    return d, float(ea.sum()), contract_rows(d, beta.numpy())


def test_c3_batch_sample_vs_oracle(oracle_mod, c3_structs):
    structs = c3_structs
    b = bt.StructureBatch(structs, 4.9)
    d = bt.BatchedSNAP(b, W, lmax=4, rcut=4.9)
    model = potential.NNPotential(EL, d.nfeat, (30, 30), seed=0).cuda()
    sp = potential.species_index(b.symbols, EL, "cuda")
    _print("c3 batch", b.nstruct, b.natoms, b.nl.nimg)
    o = bt.energy_forces_virial(d, model, sp)
    x = o["x"].detach().cpu().numpy()
    e = o["energy"].detach().cpu().numpy()
    f = o["forces"].detach().cpu().numpy()
    a0 = b.builder.atom_start_host
    counts = np.array([len(s) for s in structs])
    rng = np.random.default_rng(3)
    sample = list(np.argsort(counts, kind="stable")[:10]) + list(rng.choice(len(structs), 10, replace=False))
    mcpu = potential.NNPotential(EL, d.nfeat, (30, 30), seed=0)
    for k in sample:
        s = structs[k]
        lo, hi = int(a0[k]), int(a0[k + 1])
        ref, e_ref, f_ref = _oracle_e_f(oracle_mod, s, mcpu)
        _print("c3 structure", k, len(s))
        assert rel_err(x[lo:hi], ref["x"]) < 1e-10
        assert abs(e[k] - e_ref) < 1e-10 * abs(e_ref)
        assert rel_err(f[lo:hi], f_ref) < 1e-9


@pytest.fixture(scope="module")
def c4_setup(c3_structs):
    structs = c3_structs[:256]
    b = bt.StructureBatch(structs, 4.9)
    d = bt.BatchedSNAP(b, W, lmax=4, rcut=4.9)
    sp = potential.species_index(b.symbols, EL, "cuda")
    teacher = potential.NNPotential(EL, d.nfeat, (30, 30), seed=100).cuda()
    labels = training.teacher_labels(d, teacher, sp, plain=True)
    _print("c4 batch", b.nstruct, b.natoms, b.nl.nimg)
    return structs, b, d, sp, teacher, labels


def test_c4_teacher_labels_equal_oracle(oracle_mod, c4_setup):
    structs, b, d, sp, teacher, labels = c4_setup
    E, F, S = (t.cpu().numpy() for t in labels)
    a0 = b.builder.atom_start_host
    tc = potential.NNPotential(EL, d.nfeat, (30, 30), seed=100)
    for k in (0, 101, 255):
        s = structs[k]
        ref, e_ref, f_ref = _oracle_e_f(oracle_mod, s, tc)
        spc = potential.species_index(s.symbols, EL, "cpu")
        x = torch.as_tensor(ref["x"]).requires_grad_(True)
        (beta,) = torch.autograd.grad(tc.atom_energies(x, spc).sum(), [x])
        s_ref = -np.einsum("ik,ikl->l", beta.numpy(), oracle_mod.rdxdr_per_atom(ref)) * training.EV2GPA
        lo, hi = int(a0[k]), int(a0[k + 1])
        assert abs(E[k] - e_ref) < 1e-10 * abs(e_ref)
        assert rel_err(F[lo:hi], f_ref) < 1e-9
        assert rel_err(S[k], s_ref) < 1e-9


def test_c4_graph_replays_and_loss_gradient(c4_setup):
    structs, b, d, sp, teacher, labels = c4_setup
    # captured step == eager step (same device-count Adam), several replays
    # with eager work and synchronisations between them
    st_e = potential.NNPotential(EL, d.nfeat, (30, 30), seed=5).cuda()
    st_g = potential.NNPotential(EL, d.nfeat, (30, 30), seed=5).cuda()
    tr_e = training.ForceTrainer(d, st_e, sp, labels, lr=1e-3, graph=False, capturable=True)
    tr_g = training.ForceTrainer(d, st_g, sp, labels, lr=1e-3, graph=True)
    le = [float(tr_e.step()) for _ in range(6)]
    lg = []
    for _ in range(4):
        lg.append(float(tr_g.step()))
        torch.cuda.synchronize()
        _ = (torch.ones(4096, device="cuda") * 3).sum().item()
    _print("c4 losses", le, lg)
    for k in range(4):
        assert abs(lg[k] - le[k + 2]) <= 1e-9 * abs(le[k + 2]), (k, lg, le)
    for pe, pg in zip(st_e.parameters(), st_g.parameters()):
        assert torch.allclose(pe, pg, rtol=1e-9, atol=1e-14)
    # d loss / d parameter = central difference of the loss
    tr = training.ForceTrainer(d, st_e, sp, labels, lr=1e-3)
    params = list(st_e.parameters())
    tr.opt.zero_grad()
    loss, _ = tr.loss()
    loss.backward()
    grads = [p.grad.clone() for p in params]
    for pi, idx in [(0, (3, 7)), (2, (5, 11)), (4, (0, 4))]:
        p = params[pi]
        v = float(p[idx].detach())
        h = 1e-6 * max(1.0, abs(v))
        with torch.no_grad():
            p[idx] = v + h
        lp = float(tr.loss()[0].detach())
        with torch.no_grad():
            p[idx] = v - h
        lm = float(tr.loss()[0].detach())
        with torch.no_grad():
            p[idx] = v
        fd = (lp - lm) / (2 * h)
        g = float(grads[pi][idx])
        assert abs(fd - g) < 1e-6 * max(abs(g), 1e-8), (pi, idx, fd, g)


def test_c5_small_cell_vs_oracle(oracle_mod):
    """SO3 6/6, rc 5, alpha 2 on a 300-atom amorphous SiO2 cell at the
    benchmark density: x, dxdr, rdxdr and seq against the oracle."""
    from oracle import so3_oracle
    s = workloads.amorphous_sio2(100, 200, seed=1)
    g = SO3(nmax=6, lmax=6, rcut=5.0, alpha=2.0, derivative=True, stress=True).calculate(s)
    r = so3_oracle.calculate(s.positions, s.cell, s.pbc, s.numbers, nmax=6, lmax=6, rcut=5.0, alpha=2.0,
                             derivative=True, stress=True)
    _print("c5 small", len(s), r["dxdr"].shape)
    check_layout(g, r)


def test_c5_full_cell(oracle_mod):
    """The 12,000-atom cell: x of every 40th centre against the oracle;
    adjoint forces against plain dP/dr forces; total force zero."""
    from oracle import so3_oracle
    from pyxtal_ff_amd import so3 as dso3
    from pyxtal_ff_amd import so4 as dso4
    s = workloads.amorphous_sio2(4000, 8000, seed=0)
    b = bt.StructureBatch([s], 5.0)
    d = bt.BatchedSO3(b, nmax=6, lmax=6, rcut=5.0, alpha=2.0)
    nl = b.nl
    _print("c5 full", b.natoms, nl.nimg)
    model = potential.NNPotential(EL, d.nfeat, (30, 30), seed=0).cuda()
    sp = potential.species_index(b.symbols, EL, "cuda")
    o = bt.energy_forces_virial(d, model, sp, virial=False)
    x = o["x"].detach().cpu().numpy()
    centres = np.arange(0, 12000, 40)
    ref = so3_oracle.calculate(s.positions, s.cell, s.pbc, s.numbers, nmax=6, lmax=6, rcut=5.0, alpha=2.0,
                               derivative=False, centres=centres)
    assert rel_err(x[centres], ref["x"]) < 1e-10
    # plain materialised dP/dr path
    xp, ct, rad = dso3.forward(nl, d.cfg, d.state.atom_z, radial=True)
    dx, _ = dso4.rows(nl, d.cfg.ncoef, dso3.dpdr(nl, d.cfg, d.state.atom_z, ct, radial=rad), stress=False)
    xx = xp.detach().requires_grad_(True)
    (beta,) = torch.autograd.grad(model.atom_energies(xx, sp).sum(), [xx])
    fp = -dso4.dxdr_contract(nl, d.cfg.ncoef, beta, dx)
    fa = o["forces"].detach()
    assert rel_err(fa.cpu().numpy(), fp.cpu().numpy()) < 1e-9
    assert float(torch.linalg.norm(fa.sum(0))) < 1e-9 * float(torch.linalg.norm(fa))
    # the whole E+F step as one HIP graph (neighbour list, expansion, power
    # spectrum, network, adjoint sweep, gather): bit for bit the eager fused
    # step, and the autograd path's forces
    from pyxtal_ff_amd.neighbors import NeighborListBuilder
    pos = torch.as_tensor(s.positions, dtype=torch.float64, device="cuda")
    spa = potential.species_atom(b.symbols, EL, "cuda")
    builder = NeighborListBuilder([len(s)], [s.cell], [s.pbc], 5.0)
    g = potential.FusedEFGraph(builder, d.cfg, (d.state.atom_z,), model, spa, pos)
    for _ in range(3):
        got = {k: v.clone() for k, v in g(pos).items()}
        torch.cuda.synchronize()
    nl2 = builder.build(pos)
    ref = potential.energy_forces_fused(nl2, dso3.SO3State(nl2, d.cfg, d.state.atom_z), model, spa)
    assert torch.equal(got["forces"], ref["forces"]) and torch.equal(got["atom_energies"], ref["atom_energies"])
    assert rel_err(got["forces"].cpu().numpy(), fa.cpu().numpy()) < 1e-10
