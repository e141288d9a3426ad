"""GPU parity of the SO(4)/SNAP path against the CPU oracle (FP64, 1e-10
relative for descriptors and derivatives, bit-exact neighbour images and
seq), plus adjoint forces / virial / double backward."""
import numpy as np
import pytest
import torch

from pyxtal_ff_amd import adjoint, potential, workloads
from pyxtal_ff_amd import so4 as dso4
from pyxtal_ff_amd.descriptors.SNAP import SO4_Bispectrum as SNAPB
from pyxtal_ff_amd.descriptors.SO4 import SO4_Bispectrum as SO4B
from pyxtal_ff_amd.neighbors import DeviceNeighborList
from tests.helpers import max_rel_elem, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _print(*a):
    print(*a, flush=True)


def check_layout(gpu, ref, tol=TOL):
    """x, dxdr, rdxdr: norm-wise relative error and elementwise error with
    a floor of 1e-6 of the array's max magnitude."""
    for key in ("x", "dxdr", "rdxdr"):
        a, b = gpu.get(key), ref.get(key)
        if b is None:
            assert a is None, key
            continue
        assert a.shape == b.shape, (key, a.shape, b.shape)
        assert rel_err(a, b) < tol, (key, rel_err(a, b))
        floor = 1e-6 * max(np.abs(b).max(), 1e-300)
        assert max_rel_elem(a, b, floor) < 1e3 * tol, (key, max_rel_elem(a, b, floor))
    if "seq" in ref:
        assert np.array_equal(gpu["seq"], ref["seq"])


def structures_for_nlist():
    rng = np.random.default_rng(11)
    out = [workloads.diamond_si(reps=2, sigma=0.05, seed=0), workloads.nacl_cluster(1.0)]
    out += [workloads.random_sio_structure(rng, nmin=8, nmax=40) for _ in range(4)]
    # a skewed cell with atoms outside the cell
    s = workloads.diamond_si(reps=2, sigma=0.05, seed=2)
    s.cell = s.cell + np.array([[0, 0, 0], [1.7, 0, 0], [0.4, -0.9, 0]])
    s.positions = s.positions + np.array([3.0, -12.0, 25.0])
    out.append(s)
    # non-periodic along z
    s = workloads.diamond_si(reps=2, sigma=0.05, seed=4)
    s.pbc = np.array([True, True, False])
    out.append(s)
    return out


def test_nlist_bitexact_vs_oracle(oracle_mod):
    structs = structures_for_nlist()
    for k, s in enumerate(structs):
        _print("nlist case", k, len(s), s.pbc)
        nl = DeviceNeighborList([s.positions], [s.cell], [s.pbc], 4.9)
        ii, jj, TT, rr = oracle_mod.nlist(s.positions, s.cell, s.pbc, 4.9)
        gi, gj, gT, gr = [t.cpu().numpy() for t in nl.images()]
        assert np.array_equal(gi, ii) and np.array_equal(gj, jj) and np.array_equal(gT, TT)
        assert np.array_equal(gr, rr)  # bit-exact separations
        lay = oracle_mod.reference_layout(s.positions, s.cell, s.pbc, 4.9, np.ones(len(s)), np.ones(len(s)))
        assert np.array_equal(nl.seq_array().cpu().numpy(), lay["seq"])
        # reverse images are exact negations
        rev = nl.img_rev[: nl.nimg].cpu().numpy()
        assert np.array_equal(gr[rev], -gr)
        assert np.array_equal(gj[rev], gi)


def test_nlist_batched_equals_single(oracle_mod):
    structs = structures_for_nlist()
    nl = DeviceNeighborList([s.positions for s in structs], [s.cell for s in structs], [s.pbc for s in structs],
                            4.9)
    gi, gj, gT, gr = [t.cpu().numpy() for t in nl.images()]
    off = 0
    pos = 0
    for s in structs:
        ii, jj, TT, rr = oracle_mod.nlist(s.positions, s.cell, s.pbc, 4.9)
        n = len(ii)
        assert np.array_equal(gi[pos:pos + n] - off, ii)
        assert np.array_equal(gj[pos:pos + n] - off, jj)
        assert np.array_equal(gr[pos:pos + n], rr)
        pos += n
        off += len(s)
    assert pos == nl.nimg


def test_nlist_dense_centres_serial_path(oracle_mod):
    """Centres with more images than the wavefront fill's LDS list (168 >
    128: two atoms in a 2.5 A cube, rc 7) are filled by one lane; batched
    with ordinary centres (diamond, skewed non-periodic dense cell), every
    list is still bit-identical to the oracle's."""
    dense = (np.array([[0.1, 0.2, 0.05], [1.3, 1.2, 1.31]]), np.eye(3) * 2.5, np.array([True] * 3))
    skew = (dense[0], np.array([[2.5, 0, 0], [0.6, 2.4, 0], [0.3, -0.2, 2.6]]), np.array([True, True, False]))
    d = workloads.diamond_si(reps=2, sigma=0.05, seed=1)
    cases = [dense, (d.positions, d.cell, d.pbc), skew]
    nl = DeviceNeighborList([c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases], 7.0)
    gi, gj, gT, gr = [t.cpu().numpy() for t in nl.images()]
    at = off = 0
    for pos, cell, pbc in cases:
        ii, jj, TT, rr = oracle_mod.nlist(pos, cell, pbc, 7.0)
        n = len(ii)
        _print("dense nlist case", len(pos), "images", n, "max per centre", np.bincount(ii).max())
        assert np.array_equal(gi[at:at + n] - off, ii) and np.array_equal(gj[at:at + n] - off, jj)
        assert np.array_equal(gT[at:at + n], TT) and np.array_equal(gr[at:at + n], rr)
        at += n
        off += len(pos)
    assert at == nl.nimg
    assert np.bincount(gi).max() > 128
    rev = nl.img_rev[: nl.nimg].cpu().numpy()
    assert np.array_equal(gr[rev], -gr) and np.array_equal(gj[rev], gi)
    # the build's reverse images (runs (i -> j) reversed into (j -> i), no
    # search on T; runs here cross the 64-image chunks) equal the checked
    # search by (j, T) that verify() runs
    fast = nl.img_rev[: nl.nimg].clone()
    nl.verify()
    assert torch.equal(nl.img_rev[: nl.nimg], fast)


def test_nlist_reverse_runs_equal_checked_search():
    """Reverse images from the run reversal equal the checked search on a
    two-species random cell, a small cell with several images per pair and
    its self-images (runs of equal neighbour longer than one), and a batch of
    both."""
    rng = np.random.default_rng(8)
    a = workloads.random_sio_structure(rng, 40, 60)
    small = (np.array([[0.0, 0.0, 0.0], [1.1, 0.9, 1.4], [2.0, 2.3, 0.4]]), np.eye(3) * 3.1, np.array([True] * 3))
    for cases, rc in (([(a.positions, a.cell, a.pbc)], 5.0), ([small], 6.0),
                      ([(a.positions, a.cell, a.pbc), small], 5.5)):
        nl = DeviceNeighborList([c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases], rc)
        n = nl.nimg
        fast = nl.img_rev[:n].clone()
        nl.verify()
        assert torch.equal(nl.img_rev[:n], fast)
        gi, gj, gT, gr = [t.cpu().numpy() for t in nl.images()]
        rv = fast.cpu().numpy()
        assert np.array_equal(gr[rv], -gr) and np.array_equal(gj[rv], gi) and np.array_equal(rv[rv], np.arange(n))


def test_nlist_unwrapped_positions_large_translations(oracle_mod):
    """Every fifth atom carried 200 c - 150 a away (unwrapped MD coordinates):
    image translations reach |T| = 201, beyond the packed 8-bit sort key, so
    the fill compares the key fields; the list is still bit-identical."""
    s = workloads.diamond_si(reps=2, sigma=0.05, seed=5)
    pos = s.positions.copy()
    pos[::5] += 200 * s.cell[2] - 150 * s.cell[0]
    nl = DeviceNeighborList([pos], [s.cell], [s.pbc], 4.9)
    ii, jj, TT, rr = oracle_mod.nlist(pos, s.cell, s.pbc, 4.9)
    gi, gj, gT, gr = [t.cpu().numpy() for t in nl.images()]
    assert np.abs(TT).max() > 128
    assert np.array_equal(gi, ii) and np.array_equal(gj, jj) and np.array_equal(gT, TT)
    assert np.array_equal(gr, rr)
    nl.verify()


def test_nlist_multi_kernel_binning_equals_single_workgroup():
    """A batch with more bins than the single-workgroup binning takes (400
    sparse 20 A boxes, > 8192 bins) is binned by the multi-kernel path; its
    images equal those of the structures built one at a time (each binned
    by one workgroup)."""
    rng = np.random.default_rng(11)
    cell = np.eye(3) * 20.0
    pos = [rng.random((12, 3)) * 20.0 for _ in range(400)]
    nl = DeviceNeighborList(pos, [cell] * 400, [[True] * 3] * 400, 4.9)
    assert nl.builder.nbins > 8192
    gi, gj, gT, gr = [t.cpu().numpy() for t in nl.images()]
    off = at = 0
    for p in pos:
        one = DeviceNeighborList([p], [cell], [[True] * 3], 4.9)
        ii, jj, TT, rr = [t.cpu().numpy() for t in one.images()]
        n = len(ii)
        assert np.array_equal(gi[at:at + n] - off, ii) and np.array_equal(gj[at:at + n] - off, jj)
        assert np.array_equal(gT[at:at + n], TT) and np.array_equal(gr[at:at + n], rr)
        at += n
        off += len(p)
    assert at == nl.nimg


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1023, 4095, 4096, 4097, 12289, 100003])
def test_scan_i64_plain_and_capped(n):
    """The device exclusive scans against numpy's cumsum: ragged lengths
    around the wavefront (64) and tile (4096) sizes; the capped variant
    clamps every offset and out[n] to the cap and reports the unclamped
    total."""
    from pyxtal_ff_amd import _native as nat
    L = nat.lib()
    rng = np.random.default_rng(n)
    host = rng.integers(0, 60, n).astype(np.int64)
    ref = np.concatenate([[0], np.cumsum(host)]).astype(np.int64)
    d_in = torch.from_numpy(host).cuda()
    out = torch.full((n + 1,), -7, dtype=torch.int64, device="cuda")
    p = lambda t: nat.ptr(t)
    nat.check(L.pxff_scan_i64(p(d_in), n, p(out), nat.stream()), "pxff_scan_i64")
    assert np.array_equal(out.cpu().numpy(), ref)
    cap = int(ref[-1] // 2)
    total = torch.full((1,), -7, dtype=torch.int64, device="cuda")
    out.fill_(-7)
    nat.check(L.pxff_scan_i64_capped(p(d_in), n, cap, p(out), p(total), nat.stream()), "pxff_scan_i64_capped")
    assert np.array_equal(out.cpu().numpy(), np.minimum(ref, cap))
    assert int(total.item()) == int(ref[-1])


def test_nlist_batch_of_large_cells_equals_single_builds():
    """Five 4096-atom structures in one batch (20480 atoms, 2560 bins, all
    counted by one workgroup): the images equal those of each structure
    built alone, with the batch's atom offsets."""
    structs = [workloads.diamond_si(reps=8, sigma=0.08, seed=40 + k) for k in range(5)]
    nl = DeviceNeighborList([s.positions for s in structs], [s.cell for s in structs],
                            [s.pbc for s in structs], 4.9)
    assert nl.natoms == 20480
    gi, gj, gT, gr = [t.cpu().numpy() for t in nl.images()]
    off = at = 0
    for s in structs:
        one = DeviceNeighborList([s.positions], [s.cell], [s.pbc], 4.9)
        ii, jj, TT, rr = [t.cpu().numpy() for t in one.images()]
        n = len(ii)
        _print("batched large-cell structure", at, n)
        assert np.array_equal(gi[at:at + n] - off, ii) and np.array_equal(gj[at:at + n] - off, jj)
        assert np.array_equal(gT[at:at + n], TT) and np.array_equal(gr[at:at + n], rr)
        at += n
        off += len(s)
    assert at == nl.nimg
    nl.verify()


@pytest.mark.parametrize("lmax,rc", [(0, 4.9), (1, 4.9), (3, 4.9), (4, 4.9), (5, 4.9)])
def test_snap_diamond64_parity(oracle_mod, lmax, rc):
    s = workloads.diamond_si(reps=2, sigma=0.05, seed=0)
    _print("snap diamond64 lmax", lmax)
    g = SNAPB({"Si": 1.0}, lmax=lmax, rcut=rc, derivative=True, stress=True).calculate(s)
    r = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, {"Si": 1.0}, lmax=lmax, rcut=rc,
                                  derivative=True, stress=True)
    check_layout(g, r)
    assert g["x"].dtype == np.float64 and g["seq"].dtype == np.int64
    assert g["elements"] == s.symbols


def test_snap_small_cells_multispecies(oracle_mod):
    rng = np.random.default_rng(5)
    w = {"Si": 1.0, "O": 0.8}
    for k in range(3):
        s = workloads.random_sio_structure(rng, nmin=8, nmax=24)
        _print("small cell", k, len(s))
        g = SNAPB(w, lmax=4, rcut=4.9, derivative=True, stress=True).calculate(s)
        r = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, w, lmax=4, rcut=4.9,
                                      derivative=True, stress=True)
        check_layout(g, r)


@pytest.mark.parametrize("cutoff", ["cosine", "tanh", "poly1", "poly2", "poly3", "poly4", "exp"])
def test_cutoffs_and_norm(oracle_mod, cutoff):
    s = workloads.nacl_cluster(1.0)
    w = {"Na": 0.3, "Cl": 0.7}
    for norm in (False, True):
        g = SNAPB(w, lmax=2, rcut=6.0, derivative=True, stress=True, normalize_U=norm,
                  cutoff_function=cutoff).calculate(s)
        r = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, w, lmax=2, rcut=6.0, derivative=True,
                                      stress=True, normalize_U=norm, cutoff_function=cutoff)
        check_layout(g, r)


def test_so4py_variant(oracle_mod):
    s = workloads.nacl_cluster(1.0)
    g = SO4B(lmax=2, rcut=6.0, derivative=True, stress=True).calculate(s)
    r = oracle_mod.so4_calculate(s.positions, s.cell, s.pbc, s.numbers, s.symbols, lmax=2, rcut=6.0,
                                 derivative=True, stress=True)
    check_layout(g, r)


def test_no_derivative_branch(oracle_mod):
    s = workloads.diamond_si(reps=2, sigma=0.05, seed=0)
    g = SNAPB({"Si": 1.0}, lmax=3, rcut=4.9, derivative=False, stress=False).calculate(s)
    r = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, {"Si": 1.0}, lmax=3, rcut=4.9,
                                  derivative=False)
    assert g["dxdr"] is None and g["rdxdr"] is None and "seq" not in g
    assert rel_err(g["x"], r["x"]) < TOL


def test_golden_drange_pin_batched():
    """x on the GPU for the reference's 214 training structures (one batched
    launch) reproduces the printed min / max-min of the shipped model."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "si_drange.npz"))
    n = g["natoms"][: int(g["ntrain"])]
    off = np.concatenate([[0], np.cumsum(n)])
    pos = [g["positions"][off[s]:off[s + 1]] for s in range(len(n))]
    nl = DeviceNeighborList(pos, g["cells"][: len(n)], g["pbc"][: len(n)], 5.0)
    cfg = dso4.So4Config(6, 5.0, 0.99363, "cosine")
    ones = torch.ones(nl.natoms, dtype=torch.float64, device="cuda")
    b, _, _ = dso4.forward(nl, cfg, ones, ones)
    x = b.cpu().numpy()
    mins, maxs = x.min(0), x.max(0)
    assert np.max(np.abs(mins - g["snap_min"])) < 2e-13
    assert np.max(np.abs((maxs - mins) - g["snap_range"])) < 4e-13


def _engine(s, lmax=4, rc=4.9, weights=None, rfac0=0.99363):
    w = weights or {"Si": 1.0}
    nl = DeviceNeighborList([s.positions], [s.cell], [s.pbc], rc)
    cfg = dso4.So4Config(2 * lmax, rc, rfac0, "cosine")
    aw = torch.as_tensor([w[e] for e in s.symbols], dtype=torch.float64, device="cuda")
    sw = torch.ones(len(s), dtype=torch.float64, device="cuda")
    elements = sorted(set(s.symbols))
    model = potential.NNPotential(elements, cfg.nb, (30, 30), seed=3).cuda()
    species = potential.species_index(s.symbols, elements, "cuda")
    state = adjoint.SO4State(nl, cfg, aw, sw)
    return nl, cfg, aw, sw, model, species, state


@pytest.mark.parametrize("lmax", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("cutoff,norm", [("cosine", False), ("tanh", True), ("poly3", False)])
def test_utot_staged_and_coefficient_free_recursion(lmax, cutoff, norm):
    """U_tot three ways on a two-species cluster with a coincident pair and
    an isolated atom (no images): the staged records with the coefficient-free
    W recursion (pxff_so4_ui_staged), the in-kernel staging with the same
    recursion (pxff_so4_ui), and the reference's recursion with its sqrt(p/q)
    coefficients (summation mode 0).  They agree to rounding, so the rescaling
    u = sqrt(m! (l-m)! / (mb! (l-mb)!)) W (a wrong factorial or a missed flip
    sign fails it) and the staged records are pinned by the reference's
    recursion."""
    from pyxtal_ff_amd import autotune
    from pyxtal_ff_amd.structure import Structure
    rng = np.random.default_rng(40 + lmax)
    pos = rng.uniform(0.0, 9.0, size=(40, 3))
    pos[1] = pos[0]  # coincident: zero weight
    pos = np.vstack([pos, [[40.0, 40.0, 40.0]]])  # isolated: only the self term
    sym = ["Si" if k % 3 else "O" for k in range(len(pos))]
    s = Structure(pos, np.zeros((3, 3)), symbols=sym, pbc=False)
    nl = DeviceNeighborList([s.positions], [s.cell], [s.pbc], 4.5)
    cfg = dso4.So4Config(2 * lmax, 4.5, 0.99363, cutoff, normalize_u=norm)
    w = {"Si": 1.0, "O": 0.8}
    aw = torch.as_tensor([w[e] for e in sym], dtype=torch.float64, device="cuda")
    sw = torch.as_tensor([1.0 + 0.1 * (k % 4) for k in range(len(sym))], dtype=torch.float64, device="cuda")
    staged = dso4.ui(nl, cfg, aw, sw).cpu().numpy()
    inkernel = dso4.ui(nl, cfg, aw, sw, staged=False).cpu().numpy()
    old = autotune.get_knob(autotune.ACCUM_LDSADD)
    try:
        autotune.set_knob(autotune.ACCUM_LDSADD, 0)
        ref = dso4.ui(nl, cfg, aw, sw, staged=False).cpu().numpy()
    finally:
        autotune.set_knob(autotune.ACCUM_LDSADD, old)
    fl = 1e-12 * float(np.abs(ref).max())
    _print("2J", 2 * lmax, cutoff, "staged/ref", max_rel_elem(staged, ref, fl), "staged==inkernel",
           bool(np.array_equal(staged, inkernel)))
    assert rel_err(staged, ref) < 1e-14 and max_rel_elem(staged, ref, fl) < 1e-3
    assert rel_err(inkernel, ref) < 1e-14
    # the isolated atom's U_tot is w_self fc(0) u_norm on the diagonal
    n = len(sym) - 1
    h = staged.reshape(-1, cfg.nhalf, 64, 2)[n // 64, :, n % 64, :]
    assert np.count_nonzero(h) > 0 and np.all(h[:, 1] == 0.0)


@pytest.mark.parametrize("lmax", [0, 1, 2, 3, 4, 5])
def test_dbdr_centres_equals_per_image(lmax):
    """The per-centre dB/dr kernel (generator form, Z staged per level from
    the tile-major layout) against the per-image kernel over the atom-major
    Z: a two-species cell with a coincident pair, 2J = 0..10."""
    rng = np.random.default_rng(21 + lmax)
    s = workloads.random_sio_structure(rng, 70, 90)
    pos = s.positions.copy()
    pos[1] = pos[0]  # coincident images contribute zero rows
    s.positions = pos
    w = {"Si": 1.0, "O": 0.8}
    nl, cfg, aw, sw, model, species, state = _engine(s, lmax=lmax, weights=w)
    _, z, _ = dso4.forward(nl, cfg, aw, sw)
    b = dso4.dbdr_per_image(nl, cfg, aw, z)
    n = nl.nimg * cfg.nb * 3
    for atom_major in (True, False):
        a = dso4.dbdr(nl, cfg, aw, z, atom_major=atom_major)
        _print("dbdr centres 2J", 2 * lmax, "atom-major Z", atom_major, "images", nl.nimg, "rel",
               rel_err(a[:n].cpu().numpy(), b[:n].cpu().numpy()))
        assert rel_err(a[:n].cpu().numpy(), b[:n].cpu().numpy()) < 1e-12
        floor = 1e-9 * float(b[:n].abs().max())
        assert max_rel_elem(a[:n].cpu().numpy(), b[:n].cpu().numpy(), floor) < 1e-9


def test_adjoint_forces_match_plain_dxdr():
    rng = np.random.default_rng(9)
    cases = [workloads.diamond_si(reps=2, sigma=0.08, seed=1), workloads.random_sio_structure(rng, 8, 30)]
    for s in cases:
        w = {"Si": 1.0, "O": 0.8}
        nl, cfg, aw, sw, model, species, state = _engine(s, weights=w)
        a = potential.energy_forces_adjoint(nl, state, model, species, stress=True)
        p = potential.energy_forces_plain(nl, cfg, aw, sw, model, species, stress=True)
        assert abs(float(a["energy"] - p["energy"])) <= 1e-12 * abs(float(p["energy"]))
        fa, fp = a["forces"].cpu().numpy(), p["forces"].cpu().numpy()
        assert rel_err(fa, fp) < 1e-9
        # stress: S_voigt = -W/V (reference sign convention, per unit volume)
        V = s.get_volume()
        W = a["virial"].cpu().numpy()
        sv = -np.array([W[0, 0], W[1, 1], W[2, 2], W[0, 1], W[0, 2], W[1, 2]]) / V
        assert rel_err(sv, p["stress_voigt"].cpu().numpy()) < 1e-9


def test_forces_entry_is_exact_negation_of_gather():
    """pxff_so4_forces (F = -dE/dR summed with the opposite sign) equals the
    negated pxff_so4_gather bit for bit; the virial is the same."""
    s = workloads.diamond_si(reps=2, sigma=0.08, seed=3)
    nl = DeviceNeighborList([s.positions], [s.cell], [s.pbc], 4.9)
    g = torch.Generator(device="cpu").manual_seed(4)
    dedr = torch.randn(nl.nimg, 3, generator=g, dtype=torch.float64).cuda()
    grad, vir = dso4.gather(nl, dedr, virial=True)
    f, vir2 = dso4.gather(nl, dedr, virial=True, forces=True)
    assert torch.equal(f, -grad) and torch.equal(vir, vir2)
    nl.verify()  # the reverse images the gather used are complete


@pytest.mark.parametrize("natoms_e", [0, 1, 64, 1000, 4097])
def test_forces_energy_entry_equals_forces_and_segment_sum(natoms_e):
    """pxff_so4_forces_energy: its forces and virial equal pxff_so4_forces
    bit for bit, and its energy (summed by the launch's extra workgroup)
    equals the one-segment pxff_segment_sum bit for bit, for energy counts
    around the workgroup's 1024-row stride (0 gives 0.0)."""
    s = workloads.diamond_si(reps=2, sigma=0.08, seed=3)
    nl = DeviceNeighborList([s.positions], [s.cell], [s.pbc], 4.9)
    g = torch.Generator(device="cpu").manual_seed(9)
    dedr = torch.randn(nl.nimg, 3, generator=g, dtype=torch.float64).cuda()
    e_atom = torch.randn(natoms_e, generator=g, dtype=torch.float64).cuda()
    f_ref, vir_ref = dso4.gather(nl, dedr, virial=True, forces=True)
    seg = torch.tensor([0, natoms_e], dtype=torch.int64, device="cuda")
    e_ref = potential._segment_total(e_atom, seg)
    f, e, vir = dso4.forces_energy(nl, dedr, e_atom, virial=True)
    assert torch.equal(f, f_ref) and torch.equal(vir, vir_ref)
    assert torch.equal(e, e_ref), (float(e), float(e_ref))
    assert abs(float(e) - float(e_atom.sum())) <= 1e-12 * max(1.0, float(e_atom.abs().sum()))
    f2, e2, none = dso4.forces_energy(nl, dedr, e_atom)
    assert none is None and torch.equal(f2, f_ref) and torch.equal(e2, e_ref)


@pytest.mark.parametrize("lmax", [1, 3, 5])
def test_fused_step_other_orders(lmax):
    """2J = 2, 6, 10 (the benchmark is 2J = 8): the fused step (U_tot, Z with
    B rows, network kernel, Y, staged dE/dr, forces entry) equals the plain
    materialised-dxdr path's energy and forces."""
    s = workloads.diamond_si(reps=2, sigma=0.08, seed=6)
    nl, cfg, aw, sw, model, species, state = _engine(s, lmax=lmax)
    sp_atom = potential.species_atom(s.symbols, model.elements, "cuda")
    f = potential.energy_forces_fused(nl, state, model, sp_atom)
    p = potential.energy_forces_plain(nl, cfg, aw, sw, model, species)
    _print("2J", 2 * lmax, "energy", float(f["energy"]), float(p["energy"]))
    assert abs(float(f["energy"] - p["energy"])) <= 1e-12 * abs(float(p["energy"]))
    assert rel_err(f["forces"].cpu().numpy(), p["forces"].cpu().numpy()) < 1e-9


def test_adjoint_energy_central_fd():
    s = workloads.diamond_si(reps=2, sigma=0.08, seed=2)
    nl, cfg, aw, sw, model, species, state = _engine(s)
    a = potential.energy_forces_adjoint(nl, state, model, species)
    f = a["forces"].cpu().numpy()
    h = 1e-5
    for m, k in [(0, 0), (21, 1), (63, 2)]:
        es = []
        for sg in (1, -1):
            t = s.copy()
            t.positions[m, k] += sg * h
            nl2 = DeviceNeighborList([t.positions], [t.cell], [t.pbc], 4.9)
            st2 = adjoint.SO4State(nl2, cfg, aw, sw)
            b = adjoint.bispectrum(nl2.img_r[: 3 * nl2.nimg].view(-1, 3), st2)
            es.append(float(model.atom_energies(b, species).sum()))
        fd = -(es[0] - es[1]) / (2 * h)
        assert abs(fd - f[m, k]) < 1e-7 * max(1.0, abs(f[m, k]))


@pytest.mark.parametrize("lmax", [4, 5])
def test_double_backward_tangent_matches_fd(lmax):
    """The backward's backward (J . dr) equals the central difference of B
    along a random per-image displacement (2J = 8: register sums; 2J = 10:
    the tuned LDS summation mode)."""
    s = workloads.diamond_si(reps=2, sigma=0.08, seed=3)
    nl, cfg, aw, sw, model, species, state = _engine(s, lmax=lmax)
    r0 = nl.img_r[: 3 * nl.nimg].view(-1, 3).clone()
    g = torch.randn(r0.shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
    db = dso4.tangent(nl, cfg, aw, dso4.forward(nl, cfg, aw, sw)[1], g, img_r=r0)
    h = 1e-5
    bp = adjoint.bispectrum(r0 + h * g, state)
    bm = adjoint.bispectrum(r0 - h * g, state)
    fd = (bp - bm) / (2 * h)
    assert rel_err(db.cpu().numpy(), fd.cpu().numpy()) < 1e-8
    # and through autograd: d/dbeta of <dE/dr, g>
    beta = torch.randn((nl.natoms, cfg.nb), dtype=torch.float64, device="cuda").requires_grad_(True)
    z = dso4.forward(nl, cfg, aw, sw)[1]
    dedr = adjoint._DEdR.apply(beta, r0, z, state)
    (gb,) = torch.autograd.grad((dedr * g).sum(), [beta])
    assert rel_err(gb.cpu().numpy(), db.cpu().numpy()) < 1e-12


def test_force_training_step_gradients():
    """Gradient of a force loss w.r.t. NN parameters through the adjoint's
    double backward equals the one through the plain dxdr path."""
    s = workloads.diamond_si(reps=2, sigma=0.08, seed=4)
    nl, cfg, aw, sw, model, species, state = _engine(s)
    target = torch.zeros((len(s), 3), dtype=torch.float64, device="cuda")
    pos = nl.pos.detach().requires_grad_(True)
    r = adjoint.separations(nl, pos)
    b = adjoint.bispectrum(r, state)
    e = model.atom_energies(b, species).sum()
    (gpos,) = torch.autograd.grad(e, [pos], create_graph=True)
    loss = ((-gpos - target) ** 2).sum()
    ga = torch.autograd.grad(loss, list(model.parameters()), allow_unused=True)
    # plain path with a materialised dxdr
    bb, z, _ = dso4.forward(nl, cfg, aw, sw)
    dx, _ = dso4.rows(nl, cfg.nb, dso4.dbdr(nl, cfg, aw, z), stress=False)
    x = bb.detach().requires_grad_(True)
    e2 = model.atom_energies(x, species).sum()
    (beta,) = torch.autograd.grad(e2, [x], create_graph=True)
    seq = nl.seq_array()
    # F_m = -sum_rows beta[i] . dxdr[(i, m)]
    contrib = torch.einsum("rb,rbk->rk", beta[seq[:, 0]], dx)
    f = -torch.zeros((len(s), 3), dtype=torch.float64, device="cuda").index_add(0, seq[:, 1], contrib)
    loss2 = ((f - target) ** 2).sum()
    gp = torch.autograd.grad(loss2, list(model.parameters()), allow_unused=True)
    for u, v in zip(ga, gp):
        # the output bias does not enter the forces
        assert (u is None) == (v is None)
        if u is not None:
            assert rel_err(u.cpu().numpy(), v.cpu().numpy()) < 1e-9


@pytest.mark.slow
def test_diamond4096_full_parity(oracle_mod):
    """4096-atom Si at 2J=8 (the benchmark structure): full x/dxdr/rdxdr
    against the oracle."""
    s = workloads.diamond_si(reps=8, sigma=0.08, seed=0)
    _print("diamond4096 gpu")
    g = SNAPB({"Si": 1.0}, lmax=4, rcut=4.9, derivative=True, stress=True).calculate(s)
    _print("diamond4096 oracle")
    r = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, {"Si": 1.0}, lmax=4, rcut=4.9,
                                  derivative=True, stress=True)
    check_layout(g, r)


def test_accumulation_modes_agree_and_reproduce():
    """The U_tot summation modes (register sums reduced once per centre,
    per-group LDS partials, LDS atomic adds issued by one wavefront,
    partials shared by two or three groups) agree, for U_tot and the forces;
    each mode is bitwise reproducible run to run (the tangent too, which
    sums in registers at 2J <= 8 whatever the mode)."""
    from pyxtal_ff_amd import autotune
    s = workloads.diamond_si(reps=3, sigma=0.08, seed=4)
    nl, cfg, aw, sw, model, species, state = _engine(s, weights={"Si": 1.0})
    g = torch.Generator(device="cpu").manual_seed(3)
    dr = torch.randn((nl.nimg, 3), generator=g, dtype=torch.float64).cuda()
    old = autotune.get_knob(autotune.ACCUM_LDSADD)
    try:
        res = {}
        for mode in (0, 1, 2, 3, 4):
            autotune.set_knob(autotune.ACCUM_LDSADD, mode)
            runs = []
            for _ in range(2):
                b, z, _ = dso4.forward(nl, cfg, aw, sw)
                f = potential.energy_forces_adjoint(nl, state, model, species)["forces"]
                t = dso4.tangent(nl, cfg, aw, z, dr)
                runs.append((b.cpu().numpy(), f.cpu().numpy(), t.cpu().numpy()))
            _print("mode", mode, "rerun max |dB|", np.abs(runs[0][0] - runs[1][0]).max())
            for k in range(3):
                assert np.array_equal(runs[0][k], runs[1][k]), (mode, k)
            res[mode] = runs[0]
        for mode in (1, 2, 3, 4):
            assert rel_err(res[mode][0], res[0][0]) < 1e-13
            assert rel_err(res[mode][1], res[0][1]) < 1e-12
            assert rel_err(res[mode][2], res[0][2]) < 1e-12
    finally:
        autotune.set_knob(autotune.ACCUM_LDSADD, old)


@pytest.mark.parametrize("act", ["Tanh", "Sigmoid", "Softplus", "ELU", "Linear"])
def test_fused_network_path_matches_autograd(act):
    """Energies, forces and virial of the fused path (network forward and
    backward in one kernel, adjoint kernels called directly) equal the
    autograd adjoint path with the torch network (two elements, drange
    scaling)."""
    rng = np.random.default_rng(21)
    s = workloads.random_sio_structure(rng, 20, 40)
    nl, cfg, aw, sw, _, species, state = _engine(s, weights={"Si": 1.0, "O": 0.8})
    x = dso4.forward(nl, cfg, aw, sw)[0].cpu().numpy()
    dr = {e: np.stack([x.min(0) - 0.1, x.max(0) + 0.1], 1) for e in ("O", "Si")}
    model = potential.NNPotential(["O", "Si"], cfg.nb, (24, 16), seed=3, drange=dr, activation=act).cuda()
    sp_idx = potential.species_index(s.symbols, ["O", "Si"], "cuda")
    sp_atom = potential.species_atom(s.symbols, ["O", "Si"], "cuda")
    a = potential.energy_forces_adjoint(nl, state, model, sp_idx, stress=True)
    f = potential.energy_forces_fused(nl, state, model, sp_atom, virial=True)
    assert rel_err(f["atom_energies"].cpu().numpy(), a["atom_energies"].cpu().numpy()) < 1e-13
    assert rel_err(f["forces"].cpu().numpy(), a["forces"].cpu().numpy()) < 1e-12
    assert rel_err(f["virial"].cpu().numpy(), a["virial"].cpu().numpy()) < 1e-12


def test_image_capacity_overflow_rebuilds():
    """A build whose image capacity is too small is detected when the image
    count is first read: the list is redone with a larger capacity (same
    images as a normal build) and the fused E+F path restarts on it."""
    from pyxtal_ff_amd.neighbors import NeighborListBuilder
    s = workloads.diamond_si(reps=2, sigma=0.08, seed=3)
    pos = torch.as_tensor(s.positions, dtype=torch.float64, device="cuda")
    ref_b = NeighborListBuilder([len(s)], [s.cell], [s.pbc], 4.9)
    ref = ref_b.build(pos)
    small = NeighborListBuilder([len(s)], [s.cell], [s.pbc], 4.9)
    small.img_capacity = 100
    nl = small.build(pos)
    assert nl.img_capacity == 100 and nl.generation == 0
    k = nl.nimg
    assert k == ref.nimg and nl.generation == 1 and nl.img_capacity >= k
    for a, b in zip(nl.images(), ref.images()):
        assert torch.equal(a, b)
    assert torch.equal(nl.img_rev[:k], ref.img_rev[:k])
    # fused E+F on a cut list restarts and matches
    cfg = dso4.So4Config(6, 4.9, 0.99363)
    w = torch.ones(len(s), dtype=torch.float64, device="cuda")
    model = potential.NNPotential(["Si"], cfg.nb, (16,), seed=2).cuda()
    sp = potential.species_atom(s.symbols, ["Si"], "cuda")
    small.img_capacity = 100
    nl2 = small.build(pos)
    f2 = potential.energy_forces_fused(nl2, adjoint.SO4State(nl2, cfg, w, w), model, sp)
    f1 = potential.energy_forces_fused(ref, adjoint.SO4State(ref, cfg, w, w), model, sp)
    assert nl2.generation == 1
    assert torch.equal(f1["forces"], f2["forces"]) and float(f1["energy"]) == float(f2["energy"])
