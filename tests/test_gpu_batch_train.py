"""Batched structures in single launches, force-matched training through the
adjoint (finite differences of forces and of parameter gradients), and the
HIP-graph-captured training step against eager steps."""
import numpy as np
import pytest
import torch

from pyxtal_ff_amd import batch as bt
from pyxtal_ff_amd import potential, training, workloads
from pyxtal_ff_amd.descriptors.SNAP import SO4_Bispectrum as SNAPB
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

W = {"Si": 1.0, "O": 0.8}


def small_batch(n=6, seed=1, nmax=40):
    rng = np.random.default_rng(seed)
    return [workloads.random_sio_structure(rng, nmin=8, nmax=nmax) for _ in range(n)]


def test_batched_snap_equals_single():
    structs = small_batch(6)
    b = bt.StructureBatch(structs, 4.9)
    d = bt.BatchedSNAP(b, W, lmax=4, rcut=4.9)
    x = d.descriptors(b.nl.img_r[: 3 * b.nl.nimg].view(-1, 3)).cpu().numpy()
    off = 0
    for s in structs:
        g = SNAPB(W, lmax=4, rcut=4.9, derivative=False).calculate(s)
        assert rel_err(x[off:off + len(s)], g["x"]) < 1e-13
        off += len(s)


def test_batched_forces_and_virial_match_plain():
    structs = small_batch(5, seed=2)
    b = bt.StructureBatch(structs, 4.9)
    d = bt.BatchedSNAP(b, W, lmax=4, rcut=4.9)
    el = ["O", "Si"]
    model = potential.NNPotential(el, d.nfeat, (30, 30), seed=1).cuda()
    sp = potential.species_index(b.symbols, el, "cuda")
    o = bt.energy_forces_virial(d, model, sp)
    E, F, S = training.teacher_labels(d, model, sp, plain=True)
    assert rel_err(o["energy"].detach().cpu().numpy(), E.cpu().numpy()) < 1e-12
    assert rel_err(o["forces"].detach().cpu().numpy(), F.cpu().numpy()) < 1e-9
    assert rel_err(bt.stress_voigt(o["virial"], b).detach().cpu().numpy() * training.EV2GPA, S.cpu().numpy()) < 1e-9


def test_batched_so3_matches_plain():
    structs = small_batch(4, seed=3)
    b = bt.StructureBatch(structs, 5.0)
    d = bt.BatchedSO3(b, nmax=4, lmax=4, rcut=5.0)
    el = ["O", "Si"]
    model = potential.NNPotential(el, d.nfeat, (30, 30), seed=2).cuda()
    sp = potential.species_index(b.symbols, el, "cuda")
    o = bt.energy_forces_virial(d, model, sp)
    E, F, S = training.teacher_labels(d, model, sp, plain=True)
    assert rel_err(o["forces"].detach().cpu().numpy(), F.cpu().numpy()) < 1e-9
    assert rel_err(bt.stress_voigt(o["virial"], b).detach().cpu().numpy() * training.EV2GPA, S.cpu().numpy()) < 1e-9


def _trainer(structs, graph=False, seed_student=5, capturable=None, materialized=False):
    b = bt.StructureBatch(structs, 4.9)
    d = bt.BatchedSNAP(b, W, lmax=4, rcut=4.9)
    el = ["O", "Si"]
    sp = potential.species_index(b.symbols, el, "cuda")
    teacher = potential.NNPotential(el, d.nfeat, (30, 30), seed=100).cuda()
    labels = training.teacher_labels(d, teacher, sp, plain=True)
    student = potential.NNPotential(el, d.nfeat, (30, 30), seed=seed_student).cuda()
    tr = training.ForceTrainer(d, student, sp, labels, force_coefficient=0.1, stress_coefficient=0.01, lr=1e-3,
                               graph=graph, capturable=capturable, materialized=materialized)
    return tr, b, d, sp


def test_adjoint_forces_central_fd_batch():
    structs = small_batch(3, seed=4, nmax=20)
    tr, b, d, sp = _trainer(structs)
    o = bt.energy_forces_virial(d, tr.model, sp, virial=False)
    f = o["forces"].detach().cpu().numpy()
    h = 1e-5
    pos0 = b.pos.clone()
    for a, k in [(0, 0), (11, 2), (int(b.natoms) - 1, 1)]:
        es = []
        for sg in (1, -1):
            p = pos0.clone()
            p[a, k] += sg * h
            nl2 = b.builder.build(p)
            st = type(d.state)(nl2, d.cfg, d.atom_w, d.site_w)
            from pyxtal_ff_amd import adjoint
            x = adjoint.bispectrum(nl2.img_r[: 3 * nl2.nimg].view(-1, 3), st)
            es.append(float(tr.model.atom_energies(x, sp).sum().detach()))
        fd = -(es[0] - es[1]) / (2 * h)
        assert abs(fd - f[a, k]) < 1e-7 * max(1.0, abs(f[a, k]))


def test_training_gradient_central_fd():
    """d(loss)/d(parameter) through the adjoint's double backward equals a
    central difference of the loss."""
    structs = small_batch(3, seed=6, nmax=20)
    tr, b, d, sp = _trainer(structs)
    params = list(tr.model.parameters())
    tr.opt.zero_grad()
    loss, _ = tr.loss()
    loss.backward()
    grads = [p.grad.clone() for p in params]
    for pi, idx in [(0, (3, 7)), (2, (5, 11)), (4, (0, 4))]:
        p = params[pi]
        h = 1e-6 * max(1.0, abs(float(p[idx])))
        v = float(p[idx])
        with torch.no_grad():
            p[idx] = v + h
        lp = float(tr.loss()[0])
        with torch.no_grad():
            p[idx] = v - h
        lm = float(tr.loss()[0])
        with torch.no_grad():
            p[idx] = v
        fd = (lp - lm) / (2 * h)
        g = float(grads[pi][idx])
        assert abs(fd - g) < 1e-6 * max(abs(g), 1e-8), (pi, idx, fd, g)


def test_graph_training_replays_match_eager():
    """Replays of the captured step reproduce eager steps (each replay checked;
    eager work and synchronisations between replays)."""
    structs = small_batch(4, seed=7, nmax=30)
    # the same (device step count) Adam in both, so the updates agree bit for bit
    tr_e, _, _, _ = _trainer(structs, graph=False, capturable=True)
    tr_g, _, _, _ = _trainer(structs, graph=True)
    # the graph trainer runs 2 eager warm-up steps before capture
    le = [float(tr_e.step()) for _ in range(6)]
    lg = []
    for k in range(4):
        lg.append(float(tr_g.step()))
        torch.cuda.synchronize()
        _ = (torch.ones(1000, device="cuda") * 2).sum().item()  # eager work between replays
    for k in range(4):
        assert abs(lg[k] - le[k + 2]) <= 1e-9 * abs(le[k + 2]), (k, lg, le)
    assert le[-1] < le[0]


def test_materialized_training_equals_adjoint():
    """Training on the materialised dxdr / per-atom rdxdr (the reference's
    way) gives the adjoint trainer's loss, forces, stress and parameter
    gradients; its captured step replays its eager steps."""
    structs = small_batch(4, seed=8, nmax=30)
    tr_a, _, _, _ = _trainer(structs)
    tr_m, _, _, _ = _trainer(structs, materialized=True)
    Ea, Fa, Sa = tr_a.predict()
    Em, Fm, Sm = tr_m.predict()
    assert rel_err(Em.detach().cpu().numpy(), Ea.detach().cpu().numpy()) < 1e-13
    assert rel_err(Fm.detach().cpu().numpy(), Fa.detach().cpu().numpy()) < 1e-10
    assert rel_err(Sm.detach().cpu().numpy(), Sa.detach().cpu().numpy()) < 1e-10
    grads = []
    for tr in (tr_a, tr_m):
        tr.opt.zero_grad()
        loss, _ = tr.loss()
        loss.backward()
        grads.append([p.grad.clone() for p in tr.model.parameters()])
    for ga, gm in zip(*grads):
        assert rel_err(gm.cpu().numpy(), ga.cpu().numpy()) < 1e-9
    tr_e, _, _, _ = _trainer(structs, graph=False, capturable=True, materialized=True)
    tr_g, _, _, _ = _trainer(structs, graph=True, materialized=True)
    le = [float(tr_e.step()) for _ in range(5)]
    lg = []
    for _ in range(3):
        lg.append(float(tr_g.step()))
        torch.cuda.synchronize()
        _ = (torch.ones(1000, device="cuda") * 2).sum().item()
    for k in range(3):
        assert abs(lg[k] - le[k + 2]) <= 1e-9 * abs(le[k + 2]), (k, lg, le)
