"""Pins of the oracle parts that no reference-produced data covers: the
cutoff functions other than cosine (values at the end points, derivatives
against central differences, through the whole dx/dr pipeline), the
``normalize_U`` scaling, the neighbour-weight conventions of SNAP and SO3
(closed forms that fail if the centre's weight were used), and the
isolated atom (CG orthogonality gives B in closed form).

Each check is chosen so that a plausible slip -- a dropped 1/rc, the
centre's weight for the neighbour's, a normalisation applied to one factor
only, a self term missing -- makes it fail."""
import numpy as np
import pytest

from pyxtal_ff_amd import workloads
from pyxtal_ff_amd.structure import Structure
from tests.helpers import rel_err

SNAP_CUTOFFS = ["cosine", "tanh", "poly1", "poly2", "poly3", "poly4"]


@pytest.fixture(scope="module")
def so3o(oracle_mod):
    from oracle import so3_oracle
    return so3_oracle


# ------------------------------------------------------------- cutoffs
@pytest.mark.parametrize("name", SNAP_CUTOFFS + ["exp"])
def test_cutoff_end_values(oracle_mod, name):
    """fc(rc) = 0 for every cutoff and fc(0) = 1 except tanh, whose value at
    0 is tanh(1)^3 by its definition (SNAP.py:349-352); fc = 0 beyond rc."""
    rc = 4.9
    f0, _ = oracle_mod.cutoff(name, 0.0, rc)
    f1, _ = oracle_mod.cutoff(name, rc, rc)
    f2, d2 = oracle_mod.cutoff(name, rc * 1.01, rc)
    assert f0 == pytest.approx(np.tanh(1.0) ** 3 if name == "tanh" else 1.0, abs=1e-15)
    assert abs(f1) < 1e-15
    assert f2 == 0.0 and d2 == 0.0


@pytest.mark.parametrize("name", SNAP_CUTOFFS)
def test_cutoff_derivative_is_central_difference(oracle_mod, name):
    """dfc matches a central difference of fc inside (0, rc), for two rc (a
    missing 1/rc factor shows as a ratio of rc)."""
    for rc in (3.7, 5.3):
        for r in np.linspace(0.05, 0.97, 11) * rc:
            # fourth-order stencil; near rc the polynomials cancel to ~1e-14
            h = 1e-4
            f = [oracle_mod.cutoff(name, r + k * h, rc)[0] for k in (-2, -1, 1, 2)]
            fd = (f[0] - 8 * f[1] + 8 * f[2] - f[3]) / (12 * h)
            d = oracle_mod.cutoff(name, r, rc)[1]
            assert abs(fd - d) < 1e-9 * max(1.0, abs(d)), (name, rc, r)


def test_exp_cutoff_derivative_is_the_references_formula(oracle_mod):
    """SNAP.py:432-438 returns 2 x exp(1 - 1/(1-x^2)) / (1+x^2)^2 as the
    derivative of exp(1 - 1/(1-x^2)), x = r/rc: not its derivative (that is
    -2 x exp(..) / ((1-x^2)^2 rc)).  The oracle (and the device kernels) keep
    the reference's expression, so dxdr / forces with this cutoff are the
    reference's, not -grad E; the central difference disagrees."""
    rc = 4.9
    for r in (0.5, 1.7, 3.1, 4.4):
        x = r / rc
        e = np.exp(1.0 - 1.0 / (1.0 - x * x))
        f, d = oracle_mod.cutoff("exp", r, rc)
        assert f == pytest.approx(e, rel=1e-14)
        assert d == pytest.approx(2 * x * e / (1 + x * x) ** 2, rel=1e-14)
        true = -2 * x * e / ((1 - x * x) ** 2 * rc)
        fd = (oracle_mod.cutoff("exp", r + 1e-6, rc)[0] - oracle_mod.cutoff("exp", r - 1e-6, rc)[0]) / 2e-6
        assert fd == pytest.approx(true, rel=1e-7)
        assert abs(d - true) > 1e-3 * abs(true)


@pytest.mark.parametrize("name", ["tanh", "poly1", "poly2", "poly3", "poly4"])
def test_snap_dxdr_central_fd_other_cutoffs(oracle_mod, name):
    """The whole dB/dr pipeline with each cutoff against central differences
    of x (two species, self images: an 8-atom periodic cell)."""
    rng = np.random.default_rng(11)
    s = workloads.random_sio_structure(rng, nmin=8, nmax=8)
    w = {"Si": 1.0, "O": 0.8}

    def calc(t, derivative):
        return oracle_mod.snap_calculate(t.positions, t.cell, t.pbc, t.symbols, w, lmax=2, rcut=4.9,
                                         derivative=derivative, stress=False, cutoff_function=name)
    d = calc(s, True)
    dense = oracle_mod.dense_dxdr(d)
    h = 1e-5
    for m, k in [(0, 0), (3, 1), (7, 2)]:
        sp, sm = s.copy(), s.copy()
        sp.positions[m, k] += h
        sm.positions[m, k] -= h
        fd = (calc(sp, False)["x"] - calc(sm, False)["x"]) / (2 * h)
        assert rel_err(dense[:, m, :, k], fd) < 1e-7, (name, m, k)


@pytest.mark.parametrize("name", ["tanh", "poly1", "poly2", "poly3", "poly4"])
def test_so3_dxdr_central_fd_other_cutoffs(so3o, oracle_mod, name):
    s = workloads.nacl_cluster(1.0)
    kw = dict(nmax=2, lmax=2, rcut=6.0, cutoff=name)
    d = so3o.calculate(s.positions, s.cell, s.pbc, s.numbers, derivative=True, stress=False, **kw)
    dense = oracle_mod.dense_dxdr(d)
    h = 1e-5
    for m, k in [(0, 0), (2, 1), (5, 2)]:
        sp, sm = s.copy(), s.copy()
        sp.positions[m, k] += h
        sm.positions[m, k] -= h
        fd = (so3o.calculate(sp.positions, s.cell, s.pbc, s.numbers, derivative=False, **kw)["x"]
              - so3o.calculate(sm.positions, s.cell, s.pbc, s.numbers, derivative=False, **kw)["x"]) / (2 * h)
        assert rel_err(dense[:, m, :, k], fd) < 1e-7, (name, m, k)


def test_so3_cutoff_table_matches_snap(so3o, oracle_mod):
    """SO3.py:378-440 restates the SNAP cutoffs; both oracle copies agree."""
    for name in SNAP_CUTOFFS:
        fn = so3o.cutoff_fn(name)
        r = np.linspace(0.1, 4.8, 9)
        f = fn(r, 4.9)
        d = fn(r, 4.9, True)
        ref = np.array([oracle_mod.cutoff(name, v, 4.9) for v in r])
        # evaluation order (x**4 against x*x*x*x, FMA contraction in the C
        # build) moves the polynomials by ~1e-14 near rc
        assert np.allclose(f, ref[:, 0], rtol=0, atol=5e-14)
        assert np.allclose(d, ref[:, 1], rtol=0, atol=5e-14)


# ---------------------------------------------------------- normalize_U
def test_normalize_u_scales_by_three_level_factors(oracle_mod):
    """normalize_U multiplies U_tot^l (self term included) and dU^l by
    n(l) = 4 pi / sqrt(l + 1) (SNAP.py:1332-1341, 1456, 1481), so every
    B_(l1, l2, l) and its derivative scale by n(l1) n(l2) n(l)."""
    rng = np.random.default_rng(5)
    s = workloads.random_sio_structure(rng, nmin=10, nmax=10)
    w = {"Si": 1.0, "O": 0.8}
    kw = dict(lmax=3, rcut=4.9, derivative=True, stress=True)
    a = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, w, normalize_U=False, **kw)
    b = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, w, normalize_U=True, **kw)
    idxb = oracle_mod.so4_tables(6)["idxb"]
    n = 4 * np.pi / np.sqrt(idxb + 1.0)
    f = n[:, 0] * n[:, 1] * n[:, 2]
    assert rel_err(b["x"], a["x"] * f) < 1e-13
    assert rel_err(b["dxdr"], a["dxdr"] * f[None, :, None]) < 1e-13
    assert rel_err(b["rdxdr"], a["rdxdr"] * f[None, :, None, None]) < 1e-13


# ----------------------------------------------------- weight conventions
def _dimer(sym_a, sym_b, r, box=30.0):
    pos = np.array([[10.0, 10.0, 10.0], [10.0 + r * 0.6, 10.0 + r * 0.48, 10.0 + r * 0.64]])
    return Structure(pos, np.eye(3) * box, symbols=[sym_a, sym_b], pbc=False)


def test_snap_dimer_b000_uses_the_neighbour_weight(oracle_mod):
    """Dimer Si-O at r < rc, no periodic images: centre i has U_tot^0 =
    w_self + w_j fc(r), so B_000(i) = (w_self + w_j fc)^3 with the
    neighbour's weight w_j (SNAP.py:273, 1419-1434) and the self weight 1
    (SNAP.py:300-301).  With the SO4 class (SO4.py): w_self = Z_i, w_j = Z_j."""
    r = 2.3
    s = _dimer("Si", "O", r)
    w = {"Si": 1.0, "O": 0.8}
    fc = oracle_mod.cutoff("cosine", r, 4.9)[0]
    d = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, w, lmax=2, rcut=4.9, derivative=False)
    assert d["x"][0, 0] == pytest.approx((1.0 + 0.8 * fc) ** 3, rel=1e-13)
    assert d["x"][1, 0] == pytest.approx((1.0 + 1.0 * fc) ** 3, rel=1e-13)
    o = oracle_mod.so4_calculate(s.positions, s.cell, s.pbc, s.numbers, s.symbols, lmax=2, rcut=4.9,
                                 derivative=False)
    assert o["x"][0, 0] == pytest.approx((14.0 + 8.0 * fc) ** 3, rel=1e-13)
    assert o["x"][1, 0] == pytest.approx((8.0 + 14.0 * fc) ** 3, rel=1e-13)


def test_so3_dimer_uses_the_neighbour_number(so3o):
    """SO3 dimer: ctot(i) = Z_j c(r_ij) and |c(r)|^2 = |c(-r)|^2 per (n, n',
    l) (Y_lm(-r) = (-1)^l Y_lm(r)), so x(i) / x(j) = (Z_j / Z_i)^2 exactly
    (SO3.py:350-353, the weight is the neighbour's atomic number)."""
    s = _dimer("Si", "O", 2.1)
    d = so3o.calculate(s.positions, s.cell, s.pbc, s.numbers, nmax=3, lmax=3, rcut=5.0, derivative=False)
    x = d["x"]
    mask = np.abs(x[1]) > 1e-12 * np.abs(x[1]).max()
    assert np.allclose(x[0][mask] / x[1][mask], (8.0 / 14.0) ** 2, rtol=1e-12)


def test_so3_weight_on_flips_other_species_only(so3o):
    """weight_on negates the weight of neighbours of another species
    (SO3.py:350-353).  For a centre A with a same-species neighbour B and an
    other-species neighbour C, P is quadratic in ctot = Z_B c_B +- Z_C c_C:
    x_on + x_off = 2 (x_AB + x_AC) with the dimers AB and AC, x_on != x_off,
    and without another species weight_on changes nothing."""
    pA, pB, pC = np.array([10.0, 10.0, 10.0]), np.array([12.1, 10.4, 9.7]), np.array([9.2, 11.9, 10.8])
    box = np.eye(3) * 30.0

    def x_of(pos, syms, on):
        st = Structure(np.array(pos), box, symbols=syms, pbc=False)
        return so3o.calculate(st.positions, st.cell, st.pbc, st.numbers, nmax=3, lmax=2, rcut=5.0,
                              derivative=False, weight_on=on)["x"][0]
    on = x_of([pA, pB, pC], ["Si", "Si", "O"], True)
    off = x_of([pA, pB, pC], ["Si", "Si", "O"], False)
    ab = x_of([pA, pB], ["Si", "Si"], False)
    ac = x_of([pA, pC], ["Si", "O"], False)
    assert rel_err(on + off, 2 * (ab + ac)) < 1e-13
    assert rel_err(on, off) > 1e-3
    same_on = x_of([pA, pB, pC], ["Si", "Si", "Si"], True)
    same_off = x_of([pA, pB, pC], ["Si", "Si", "Si"], False)
    assert rel_err(same_on, same_off) < 1e-15


# ----------------------------------------------------------- isolated atom
@pytest.mark.parametrize("cutoff", ["cosine", "tanh"])
def test_isolated_atom_closed_form(oracle_mod, cutoff):
    """An atom with no neighbour: U_tot^l = w fc(0) I (the self term,
    SNAP.py:590-627), Z = (w fc(0))^2 I by CG orthogonality, hence
    B_(l1, l2, l) = (w fc(0))^3 (l + 1) for every component; its only seq
    row (0, 0) has zero derivative.  fc(0) = 1 for cosine, tanh(1)^3 for tanh."""
    s = Structure([[1.0, 2.0, 3.0]], np.eye(3) * 20.0, symbols=["Si"], pbc=True)
    d = oracle_mod.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, {"Si": 1.0}, lmax=4, rcut=4.9,
                                  derivative=True, stress=True, cutoff_function=cutoff)
    idxb = oracle_mod.so4_tables(8)["idxb"]
    fc0 = oracle_mod.cutoff(cutoff, 0.0, 4.9)[0]
    assert np.allclose(d["x"][0], fc0 ** 3 * (idxb[:, 2] + 1.0), rtol=1e-13, atol=0)
    assert d["seq"].tolist() == [[0, 0]]
    assert not d["dxdr"].any() and not d["rdxdr"].any()
