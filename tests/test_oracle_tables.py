"""Pin the oracle's Clebsch-Gordan table and Wigner-U recursion against
closed forms computed here independently (exact rational Racah sums; the
symmetric-power SU(2) representation)."""
import math
from fractions import Fraction

import numpy as np
import pytest


def racah_cg_squared_signed(j1, m1, j2, m2, j, m):
    """<j1 m1 j2 m2 | j m> as (sign, exact square) with all arguments doubled
    (2j etc.), by Racah's formula."""
    if m1 + m2 != m or abs(m1) > j1 or abs(m2) > j2 or abs(m) > j:
        return 0, Fraction(0)
    f = lambda x2: math.factorial(x2 // 2)  # argument is 2x, always even here
    pre = Fraction((j + 1) * f(j1 + j2 - j) * f(j1 - j2 + j) * f(-j1 + j2 + j), f(j1 + j2 + j + 2))
    pre *= f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j + m) * f(j - m)
    s = Fraction(0)
    for k2 in range(0, 2 * (j1 + j2 + j) + 2, 2):
        args = [j1 + j2 - j - k2, j1 - m1 - k2, j2 + m2 - k2, j - j2 + m1 + k2, j - j1 - m2 + k2]
        if min(args) < 0:
            continue
        den = f(k2) * f(args[0]) * f(args[1]) * f(args[2]) * f(args[3]) * f(args[4])
        s += Fraction((-1) ** (k2 // 2), den)
    if s == 0:
        return 0, Fraction(0)
    return (1 if s > 0 else -1), pre * s * s


@pytest.mark.parametrize("twol", [2, 4, 6, 8])
def test_cg_matches_racah(oracle_mod, twol):
    t = oracle_mod.so4_tables(twol)
    cg = t["cglist"]
    k = 0
    worst = 0.0
    for l1 in range(twol + 1):
        for l2 in range(l1 + 1):
            for l in range(l1 - l2, min(twol, l1 + l2) + 1, 2):
                for m1 in range(l1 + 1):
                    aa2 = 2 * m1 - l1
                    for m2 in range(l2 + 1):
                        bb2 = 2 * m2 - l2
                        sgn, sq = racah_cg_squared_signed(l1, aa2, l2, bb2, l, aa2 + bb2)
                        exact = sgn * math.sqrt(float(sq))
                        worst = max(worst, abs(cg[k] - exact))
                        k += 1
    assert k == len(cg)
    assert worst < 1e-13


def symmetric_power_u(l, a, b):
    """U^l of the SU(2) matrix M = [[conj a, -conj b], [b, a]] in the
    normalised symmetric-power basis: row mb, column ma."""
    al, be, ga, de = np.conj(a), -np.conj(b), b, a
    U = np.zeros((l + 1, l + 1), dtype=complex)
    for mb in range(l + 1):
        # coefficients of (al x + be y)^(l-mb) (ga x + de y)^mb in x^(l-ma) y^ma
        p1 = np.array([math.comb(l - mb, k) * al ** (l - mb - k) * be ** k for k in range(l - mb + 1)])
        p2 = np.array([math.comb(mb, k) * ga ** (mb - k) * de ** k for k in range(mb + 1)])
        c = np.convolve(p1, p2)
        for ma in range(l + 1):
            norm = math.sqrt(math.factorial(ma) * math.factorial(l - ma) / (math.factorial(mb) * math.factorial(l - mb)))
            U[mb, ma] = c[ma] * norm
    return U


@pytest.mark.parametrize("twol", [4, 8])
def test_u_recursion_matches_symmetric_power(oracle_mod, twol):
    rng = np.random.default_rng(1)
    for _ in range(5):
        v = rng.normal(size=3)
        r = np.linalg.norm(v)
        psi = 0.99363 * np.pi * r / 4.9
        u = oracle_mod.uarray(twol, v[0], v[1], v[2], psi, r)
        a = np.cos(psi) - 1j * np.sin(psi) * v[2] / r
        b = np.sin(psi) / r * (v[1] - 1j * v[0])
        off = 0
        for l in range(twol + 1):
            blk = u[off:off + (l + 1) ** 2].reshape(l + 1, l + 1)
            ref = symmetric_power_u(l, a, b)
            assert np.max(np.abs(blk - ref)) < 1e-13
            # unitarity
            assert np.max(np.abs(blk @ blk.conj().T - np.eye(l + 1))) < 1e-12
            off += (l + 1) ** 2


def test_du_recursion_matches_fd(oracle_mod):
    twol, rc, rf = 8, 4.9, 0.99363
    v = np.array([1.3, -0.7, 2.1])
    r = np.linalg.norm(v)
    _, du = oracle_mod.uarray(twol, *v, rf * np.pi * r / rc, r, derivative=True)
    h = 1e-6
    for k in range(3):
        vp, vm = v.copy(), v.copy()
        vp[k] += h
        vm[k] -= h
        rp, rm = np.linalg.norm(vp), np.linalg.norm(vm)
        up = oracle_mod.uarray(twol, *vp, rf * np.pi * rp / rc, rp)
        um = oracle_mod.uarray(twol, *vm, rf * np.pi * rm / rc, rm)
        fd = (up - um) / (2 * h)
        assert np.max(np.abs(du[:, k] - fd)) < 1e-8


def test_u_identity_at_zero_rotation(oracle_mod):
    u = oracle_mod.uarray(6, 0.3, 0.4, 0.5, 0.0, np.sqrt(0.5))
    off = 0
    for l in range(7):
        blk = u[off:off + (l + 1) ** 2].reshape(l + 1, l + 1)
        assert np.allclose(blk, np.eye(l + 1), atol=1e-15)
        off += (l + 1) ** 2


def test_table_sizes(oracle_mod):
    # counts of coefficients, Z elements and U elements of the reference's
    # index construction
    s6, s8 = oracle_mod.so4_sizes(6), oracle_mod.so4_sizes(8)
    assert (s6["nb"], s6["nz"], s6["nu"], s6["ncg"]) == (30, 779, 140, 1321)
    assert (s8["nb"], s8["nz"], s8["nu"], s8["ncg"]) == (55, 2386, 285, 4098)
