"""The training loss against a restatement of the reference's
``calculate_loss`` (neuralnetwork.py:466-545) on oracle descriptors: per
structure, per element, forces by the dense dxdr contraction
einsum("ik,ijkl->jl") and stress by einsum("ik,ikl->l") over the per-atom
rdxdr, exactly the reference's loop.  CPU only (the network runs in torch on
the host)."""
import numpy as np
import pytest
import torch

from pyxtal_ff_amd import potential, training, workloads

W = {"Si": 1.0, "O": 0.8}
ELEMENTS = ["O", "Si"]


def _records(oracle_mod, structs):
    out = []
    for s in structs:
        d = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, W, lmax=2, rcut=4.9,
                                      derivative=True, stress=True)
        dense = oracle_mod.dense_dxdr(d)  # [N, N, nc, 3]
        rv = oracle_mod.rdxdr_per_atom(d)  # [N, nc, 6]
        sym = np.array(s.symbols)
        rec = {"n": len(s), "x": {}, "dxdr": {}, "rdxdr": {}}
        for e in ELEMENTS:
            idx = np.nonzero(sym == e)[0]
            rec["x"][e] = torch.tensor(d["x"][idx])
            rec["dxdr"][e] = torch.tensor(dense[idx])
            rec["rdxdr"][e] = torch.tensor(rv[idx])
        out.append(rec)
    return out


def _reference_calculate_loss(model, recs, labels, sf, cF, cS, stress_groups, groups, alpha):
    """neuralnetwork.py:466-545, restated statement by statement."""
    energy_loss = force_loss = stress_loss = 0.0
    all_atoms = 0
    s_count = 0
    preds = []
    for rec, (energy, force, stress), w, group in zip(recs, labels, sf, groups):
        n_atoms = rec["n"]
        all_atoms += n_atoms
        _Energy = 0
        _force = torch.zeros([n_atoms, 3], dtype=torch.float64)
        use_stress = cS and (group in stress_groups)
        if use_stress:
            _stress = torch.zeros([6], dtype=torch.float64)
        for element in ELEMENTS:
            x = rec["x"][element]
            if x.nelement() > 0:
                _x = x.clone().requires_grad_()
                _energy = model.models[element](_x).sum()
                _Energy += _energy
                dedx = torch.autograd.grad(_energy, _x, create_graph=True)[0]
                if cF:
                    _force -= torch.einsum("ik, ijkl->jl", dedx, rec["dxdr"][element])
                if use_stress:
                    _stress += -1 * torch.einsum("ik, ikl->l", dedx, rec["rdxdr"][element])
        energy_loss += w * ((_Energy - energy) / n_atoms) ** 2
        if cF:
            force_loss += w * cF * ((_force - force) ** 2).sum()
        if use_stress:
            _stress = _stress * training.EV2GPA
            stress_loss += w * cS * ((_stress - stress) ** 2).sum()
            s_count += 6
        preds.append((_Energy, _force, _stress if use_stress else torch.zeros(6, dtype=torch.float64)))
    energy_loss = energy_loss / (2.0 * len(recs))
    loss = energy_loss
    if cF:
        loss = loss + force_loss / (2.0 * all_atoms)
    if cS:
        loss = loss + stress_loss / (2.0 * s_count)
    if alpha:
        reg = 0.0
        for element, m in model.models.items():
            for name, p in m.named_parameters():
                if "weight" in name:
                    reg += alpha * p.pow(2).sum()
        loss = loss + reg
    return loss, preds


@pytest.mark.parametrize("cF,cS,alpha,mixed_groups", [(0.1, 0.01, None, False), (0.3, 0.02, 1e-3, True),
                                                      (0.1, None, None, False)])
def test_reference_loss_equals_calculate_loss(oracle_mod, cF, cS, alpha, mixed_groups):
    rng = np.random.default_rng(21)
    structs = [workloads.random_sio_structure(rng, nmin=8, nmax=14) for _ in range(3)]
    recs = _records(oracle_mod, structs)
    nf = recs[0]["x"]["Si"].shape[1]
    model = potential.NNPotential(ELEMENTS, nf, (6, 5), seed=3)
    g = torch.Generator().manual_seed(4)
    labels = [(torch.randn((), generator=g, dtype=torch.float64) * 3 - 20,
               torch.randn((r["n"], 3), generator=g, dtype=torch.float64),
               torch.randn(6, generator=g, dtype=torch.float64) * 5) for r in recs]
    sf = [0.7, 1.0, 1.3]
    groups = ["a", "b", "a"] if mixed_groups else ["a"] * 3
    stress_groups = ["a"]
    ref, preds = _reference_calculate_loss(model, recs, labels, sf, cF, cS, stress_groups, groups, alpha)

    E = torch.stack([p[0] for p in preds])
    S = torch.stack([p[2] for p in preds])
    f_sq = torch.stack([((p[1] - lab[1]) ** 2).sum() for p, lab in zip(preds, labels)])
    E_ref = torch.stack([lab[0] for lab in labels])
    S_ref = torch.stack([lab[2] for lab in labels])
    n = torch.tensor([r["n"] for r in recs], dtype=torch.float64)
    mask = torch.tensor([1.0 if gg in stress_groups else 0.0 for gg in groups], dtype=torch.float64)
    got = training.reference_loss(E, E_ref, n, f_sq, torch.tensor(sf, dtype=torch.float64), cF, S, S_ref, mask,
                                  cS, training.weight_penalty(model, alpha))
    assert float(got) == pytest.approx(float(ref), rel=1e-13)
    params = list(model.parameters())
    g_ref = torch.autograd.grad(ref, params, retain_graph=True, allow_unused=True)
    g_got = torch.autograd.grad(got, params, allow_unused=True)
    for a, b in zip(g_ref, g_got):
        if a is None:
            assert b is None or not b.any()
            continue
        assert torch.allclose(a, b, rtol=1e-11, atol=1e-15)


def test_loss_normalisations_differ_from_per_structure_means():
    """A per-structure 1/(3n) force mean and eV/A^3 stress (the obvious
    alternatives) give other values: the test above would see them."""
    E = torch.tensor([1.0, 2.0], dtype=torch.float64)
    n = torch.tensor([4.0, 8.0], dtype=torch.float64)
    f_sq = torch.tensor([3.0, 5.0], dtype=torch.float64)
    one = torch.ones(2, dtype=torch.float64)
    ref = training.reference_loss(E, 0 * E, n, f_sq, one, 1.0)
    alt = ((E / n) ** 2).mean() + (f_sq / (3 * n)).mean()
    assert abs(float(ref) - float(alt)) > 1e-3
    assert float(ref) == pytest.approx(((1 / 4) ** 2 + (2 / 8) ** 2) / 4 + 8.0 / 24)
