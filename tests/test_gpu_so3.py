"""GPU parity of the SO(3) power spectrum against the CPU oracle (FP64,
1e-10 relative), the shipped-model drange pin on the GPU, and the adjoint
force / double-backward checks."""
import os

import numpy as np
import pytest
import torch

from pyxtal_ff_amd import so3 as dso3
from pyxtal_ff_amd import workloads, adjoint
from pyxtal_ff_amd.descriptors.SO3 import SO3
from pyxtal_ff_amd.neighbors import DeviceNeighborList
from tests.helpers import rel_err
from tests.test_gpu_so4 import check_layout

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def so3o(oracle_mod):
    from oracle import so3_oracle
    return so3_oracle


@pytest.mark.parametrize("cutoff", ["cosine", "tanh", "poly1", "poly2", "poly3", "poly4", "unity"])
def test_so3_nacl_parity_cutoffs(so3o, cutoff):
    s = workloads.nacl_cluster(1.0)
    print("so3 nacl", cutoff, flush=True)
    g = SO3(nmax=2, lmax=2, rcut=6.0, derivative=True, stress=True, cutoff_function=cutoff).calculate(s)
    r = so3o.calculate(s.positions, s.cell, s.pbc, s.numbers, nmax=2, lmax=2, rcut=6.0, derivative=True,
                       stress=True, cutoff=cutoff)
    check_layout(g, r)


@pytest.mark.parametrize("lmax", [6, 7])
def test_so3_images_on_the_polar_axis(so3o, lmax):
    """The unrotated rocksalt cluster has images straight along +-z (sin
    theta = 0, where Y_lm vanishes for m != 0 and phi is undefined): x, dxdr
    and rdxdr against the oracle through the per-(l, m) expansion sums."""
    s = workloads.nacl_cluster(0.0)
    d = s.positions[:, None, :] - s.positions[None, :, :]
    assert np.any((np.abs(d[..., 0]) < 1e-12) & (np.abs(d[..., 1]) < 1e-12) & (np.abs(d[..., 2]) > 0))
    print("so3 nacl on axis lmax", lmax, flush=True)
    g = SO3(nmax=3, lmax=lmax, rcut=6.0, derivative=True, stress=True).calculate(s)
    r = so3o.calculate(s.positions, s.cell, s.pbc, s.numbers, nmax=3, lmax=lmax, rcut=6.0, derivative=True,
                       stress=True)
    check_layout(g, r)


@pytest.mark.parametrize("nmax,lmax", [(3, 4), (6, 6), (1, 0), (4, 9)])
def test_so3_diamond_parity(so3o, nmax, lmax):
    s = workloads.diamond_si(reps=2, sigma=0.05, seed=0)
    print("so3 diamond", nmax, lmax, flush=True)
    g = SO3(nmax=nmax, lmax=lmax, rcut=5.0, alpha=2.0, derivative=True, stress=True).calculate(s)
    r = so3o.calculate(s.positions, s.cell, s.pbc, s.numbers, nmax=nmax, lmax=lmax, rcut=5.0, alpha=2.0,
                       derivative=True, stress=True)
    check_layout(g, r)


def test_so3_small_cells_weight_on(so3o):
    rng = np.random.default_rng(2)
    for k in range(3):
        s = workloads.random_sio_structure(rng, nmin=8, nmax=20)
        print("so3 small", k, len(s), flush=True)
        g = SO3(nmax=3, lmax=3, rcut=4.9, derivative=True, stress=True, weight_on=True).calculate(s)
        r = so3o.calculate(s.positions, s.cell, s.pbc, s.numbers, nmax=3, lmax=3, rcut=4.9, derivative=True,
                           stress=True, weight_on=True)
        check_layout(g, r)


def test_so3_golden_drange_pin_batched():
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "si_drange.npz"))
    n = g["natoms"][: int(g["ntrain"])]
    off = np.concatenate([[0], np.cumsum(n)])
    nl = DeviceNeighborList([g["positions"][off[s]:off[s + 1]] for s in range(len(n))], g["cells"][: len(n)],
                            g["pbc"][: len(n)], 5.0)
    cfg = dso3.So3Config(3, 4, 5.0, 2.0)
    z = torch.as_tensor(g["numbers"][: off[len(n)]], dtype=torch.int64, device="cuda")
    x, _ = dso3.forward(nl, cfg, z)
    x = x.cpu().numpy()
    mins, maxs = x.min(0), x.max(0)
    scale = np.abs(g["so3_range"]).max()
    assert np.max(np.abs(mins - g["so3_min"])) < 1e-13 * scale
    assert np.max(np.abs((maxs - mins) - g["so3_range"])) < 2e-13 * scale


def _so3_engine(s, nmax=4, lmax=4, rc=5.0):
    from pyxtal_ff_amd import potential
    nl = DeviceNeighborList([s.positions], [s.cell], [s.pbc], rc)
    cfg = dso3.So3Config(nmax, lmax, rc, 2.0)
    z = torch.as_tensor(s.numbers, dtype=torch.int64, device="cuda")
    elements = sorted(set(s.symbols))
    model = potential.NNPotential(elements, cfg.ncoef, (30, 30), seed=5).cuda()
    species = potential.species_index(s.symbols, elements, "cuda")
    return nl, cfg, z, model, species, dso3.SO3State(nl, cfg, z)


def _plain_forces(nl, cfg, z, model, species):
    from pyxtal_ff_amd import so4 as dso4
    x, ct = dso3.forward(nl, cfg, z)
    dp = dso3.dpdr(nl, cfg, z, ct)
    dx, _ = dso4.rows(nl, cfg.ncoef, dp, stress=False)
    xx = x.detach().requires_grad_(True)
    e = model.atom_energies(xx, species).sum()
    (beta,) = torch.autograd.grad(e, [xx])
    return e.detach(), -dso4.dxdr_contract(nl, cfg.ncoef, beta, dx)


def _adjoint_forces(nl, state, model, species):
    pos = nl.pos.detach().requires_grad_(True)
    r = adjoint.separations(nl, pos)
    x = dso3.power_spectrum(r, state)
    e = model.atom_energies(x, species).sum()
    (g,) = torch.autograd.grad(e, [pos])
    return e.detach(), -g


@pytest.mark.parametrize("nmax,lmax", [(4, 4), (2, 0), (3, 1), (5, 6)])
def test_so3_adjoint_matches_plain(nmax, lmax):
    """Adjoint energies and forces (per-lane Y rows to lmax + 1, lanes past
    them idle) against the plain dP/dr path."""
    rng = np.random.default_rng(4)
    for s in [workloads.diamond_si(reps=2, sigma=0.08, seed=1), workloads.random_sio_structure(rng, 8, 30)]:
        print("so3 adjoint vs plain", nmax, lmax, len(s), flush=True)
        nl, cfg, z, model, species, st = _so3_engine(s, nmax=nmax, lmax=lmax)
        ea, fa = _adjoint_forces(nl, st, model, species)
        ep, fp = _plain_forces(nl, cfg, z, model, species)
        assert abs(float(ea - ep)) <= 1e-12 * abs(float(ep))
        assert rel_err(fa.cpu().numpy(), fp.cpu().numpy()) < 1e-9


def test_so3_stored_radial_equals_recomputed():
    """Adjoint forces with the forward's stored radial integrals (all orders
    of an image from one Bessel evaluation per node) equal those with the
    integrals recomputed per (image, l) in the backward (same quadrature;
    the two Bessel evaluations agree to ~1e-15, the forces, sums with
    cancellation, to ~1e-12)."""
    rng = np.random.default_rng(5)
    s = workloads.random_sio_structure(rng, 20, 40)
    nl, cfg, z, model, species, st = _so3_engine(s)
    out = []
    for store in (True, False):
        st2 = dso3.SO3State(nl, cfg, z, store_radial=store)
        out.append(_adjoint_forces(nl, st2, model, species)[1].cpu().numpy())
    assert rel_err(out[0], out[1]) < 1e-11


@pytest.mark.parametrize("nmax,lmax", [(6, 6), (3, 2), (1, 0)])
def test_so3_radial_integrals_vs_scipy_quadrature(nmax, lmax):
    """The stored radial integrals I_nl(R) = sum_q G[n][q] i_l(2 alpha R r_q)
    and dI_nl/dR against the same quadrature with scipy.special.spherical_in
    on the host (amorphous SiO2 images, 2 alpha R r_q from ~0.001 to 100:
    the series, downward-recurrence and upward-recurrence ranges), both as
    tabulated (the default for these sizes) and by the device quadrature;
    the quadrature once per image pair equals it for every image to rounding
    (the same sums; a multiply-add may be fused in one kernel and not in the
    other)."""
    from scipy.special import spherical_in
    s = workloads.amorphous_sio2(40, 80, seed=1)
    nl = DeviceNeighborList([s.positions], [s.cell], [s.pbc], 5.0)
    cfg = dso3.So3Config(nmax, lmax, 5.0, 2.0)
    z = torch.as_tensor(s.numbers, dtype=torch.int64, device="cuda")
    n = nl.nimg
    m = n * (lmax + 1) * 2 * nmax
    # once per image pair (the list's own separations): table and quadrature
    ct, rad = dso3.ctot(nl, cfg, z, radial=True)
    ctq, radq = dso3.ctot(nl, cfg, z, radial=True, use_table=False)
    # every image evaluated by the quadrature (a copy of the separations)
    ct2, rad2 = dso3.ctot(nl, cfg, z, img_r=nl.img_r.clone(), radial=True)
    # pairs' rows: both images of a pair carry the one evaluation
    rev = nl.img_rev[:n].cpu().numpy()
    rq = radq[:m].view(n, -1).cpu().numpy()
    assert np.array_equal(rq, rq[rev])
    print("pair vs per-image quadrature", rel_err(radq[:m].cpu().numpy(), rad2[:m].cpu().numpy()),
          rel_err(ctq.cpu().numpy(), ct2.cpu().numpy()), flush=True)
    assert rel_err(radq[:m].cpu().numpy(), rad2[:m].cpu().numpy()) < 1e-14
    assert rel_err(ctq.cpu().numpy(), ct2.cpu().numpy()) < 1e-14
    print("table vs quadrature", rel_err(rad[:m].cpu().numpy(), radq[:m].cpu().numpy()), flush=True)
    assert rel_err(rad[:m].cpu().numpy(), radq[:m].cpu().numpy()) < 1e-13
    assert rel_err(ct.cpu().numpy(), ctq.cpu().numpy()) < 1e-13
    rt = rad[:m].view(n, -1).cpu().numpy()
    assert np.array_equal(rt, rt[rev])
    for store in (rad, radq):
        _check_radial_store(store[:m].view(n, lmax + 1, 2, nmax).cpu().numpy(), nl, n, nmax, lmax)


def _check_radial_store(rad, nl, n, nmax, lmax):
    from scipy.special import spherical_in
    G, Q, _ = dso3.radial_tables(nmax, lmax, 5.0, 2.0)
    r = nl.img_r[: 3 * n].view(n, 3).cpu().numpy()
    pick = np.linspace(0, n - 1, 300).astype(int)
    R = np.linalg.norm(r[pick], axis=1)
    x = 4.0 * R[:, None] * Q[None, :]  # 2 alpha R r_q
    print("x range", x.min(), x.max(), flush=True)
    assert x.min() < 0.1 and x.max() > 20.0
    for l in range(lmax + 1):
        il = spherical_in(l, x)
        dil = spherical_in(l, x, derivative=True)
        I = il @ G.T
        dI = (4.0 * Q[None, :] * dil) @ G.T
        assert rel_err(rad[pick, l, 0, :], I) < 1e-12, l
        assert rel_err(rad[pick, l, 1, :], dI) < 1e-12, l


def test_so3_double_backward_tangent_fd():
    s = workloads.diamond_si(reps=2, sigma=0.08, seed=3)
    nl, cfg, z, model, species, st = _so3_engine(s)
    r0 = nl.img_r[: 3 * nl.nimg].view(-1, 3).clone()
    g = torch.randn(r0.shape, dtype=torch.float64, device="cuda")
    x, ct = dso3.forward(nl, cfg, z, img_r=r0)
    dct = dso3.ctot(nl, cfg, z, img_r=r0, dr=g)
    dx = dso3.power(nl, cfg, ct, dct)
    h = 1e-5
    fd = (dso3.power_spectrum(r0 + h * g, st) - dso3.power_spectrum(r0 - h * g, st)) / (2 * h)
    assert rel_err(dx.cpu().numpy(), fd.cpu().numpy()) < 1e-8


@pytest.mark.slow
def test_so3_sio2_scale_adjoint_vs_plain():
    """Amorphous SiO2 (1200 atoms here; the benchmark scale uses 12000) at
    nmax = lmax = 6, rc 5.0: adjoint forces equal plain-dP/dr forces; force
    sum vanishes."""
    s = workloads.amorphous_sio2(400, 800, seed=0)
    nl, cfg, z, model, species, st = _so3_engine(s, nmax=6, lmax=6, rc=5.0)
    ea, fa = _adjoint_forces(nl, st, model, species)
    ep, fp = _plain_forces(nl, cfg, z, model, species)
    assert rel_err(fa.cpu().numpy(), fp.cpu().numpy()) < 1e-9
    assert float(torch.linalg.norm(fa.sum(0))) < 1e-9 * float(torch.linalg.norm(fa))


@pytest.mark.parametrize("nmax,lmax,rc,alpha", [(6, 6, 6.0, 4.0), (3, 3, 4.0, 0.5), (6, 4, 5.0, 6.0)])
def test_so3_radial_table_other_parameters(nmax, lmax, rc, alpha):
    """The tabulated integrals against the device quadrature for steeper and
    flatter settings (the interval count grows with 2 alpha rc^2): norm-wise
    1e-13, and every value within 1e-12 of its function's largest value."""
    s = workloads.amorphous_sio2(40, 80, seed=2)
    nl = DeviceNeighborList([s.positions], [s.cell], [s.pbc], rc)
    cfg = dso3.So3Config(nmax, lmax, rc, alpha)
    z = torch.as_tensor(s.numbers, dtype=torch.int64, device="cuda")
    m = nl.nimg * (lmax + 1) * 2 * nmax
    _, rad = dso3.ctot(nl, cfg, z, radial=True)
    _, radq = dso3.ctot(nl, cfg, z, radial=True, use_table=False)
    a = rad[:m].view(nl.nimg, -1).cpu().numpy()
    b = radq[:m].view(nl.nimg, -1).cpu().numpy()
    print("table vs quadrature", rel_err(a, b), flush=True)
    assert rel_err(a, b) < 1e-13
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    assert (np.abs(a - b) / scale).max() < 1e-12
