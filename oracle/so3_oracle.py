"""CPU oracle of the SO(3) power spectrum (TEST INFRASTRUCTURE ONLY).

NumPy/SciPy restatement of ``pyxtal_ff/descriptors/SO3.py``:
  * ``W``                      SO3.py:466-479 (S^-1/2 via inv + eig)
  * ``phi`` / ``g``            SO3.py:481-493
  * ``GaussChebyshevQuadrature`` SO3.py:495-502
  * ``compute_cs`` / ``compute_dcs`` SO3.py:504-694 (scipy ``spherical_in``
    and ``sph_harm`` exactly as the reference calls them, covariant-component
    gradient of Y_lm)
  * ``calculate``              SO3.py:187-291 (weights, norm_l, tril(n>=n')
    with l fastest, dP = sum_m dc conj(ctot) + conj-transpose, row scatter,
    pstress)
The per-centre grouping uses index arithmetic instead of the reference's
O(N^2) boolean masks (same sums, same order of rows), and images are
processed in chunks so that large structures fit in memory.
"""
import numpy as np
from scipy.special import spherical_in

try:
    # SciPy >= 1.15: sph_harm_y(l, m, polar, azimuth) is the same function as
    # the deprecated sph_harm(m, l, azimuth, polar) the reference calls
    from scipy.special import sph_harm_y as _sph_harm_y

    def sph_harm(m, l, phi, theta):
        return _sph_harm_y(l, m, theta, phi)
except ImportError:  # pragma: no cover - older SciPy
    from scipy.special import sph_harm

import oracle as _oracle


def W(nmax):
    arr = np.zeros((nmax, nmax), np.float64)
    for alpha in range(1, nmax + 1):
        temp1 = (2 * alpha + 5) * (2 * alpha + 6) * (2 * alpha + 7)
        for beta in range(1, alpha + 1):
            temp2 = (2 * beta + 5) * (2 * beta + 6) * (2 * beta + 7)
            arr[alpha - 1, beta - 1] = np.sqrt(temp1 * temp2) / (5 + alpha + beta) / (6 + alpha + beta) / (
                7 + alpha + beta)
            arr[beta - 1, alpha - 1] = arr[alpha - 1, beta - 1]
    sinv = np.linalg.inv(arr)
    eigvals, V = np.linalg.eig(sinv)
    sqrtD = np.diag(np.sqrt(eigvals))
    arr[:, :] = np.dot(np.dot(V, sqrtD), np.linalg.inv(V))
    return arr


def phi(r, alpha, rcut):
    return (rcut - r) ** (alpha + 2) / np.sqrt(2 * rcut ** (2 * alpha + 7) / (2 * alpha + 5) / (2 * alpha + 6) / (
        2 * alpha + 7))


def g(r, n, nmax, rcut, w):
    s = 0.0
    for alpha in range(1, nmax + 1):
        s += w[n - 1, alpha - 1] * phi(r, alpha, rcut)
    return s


def gauss_chebyshev(nmax, lmax):
    nq = (nmax + lmax + 1) * 10
    x = np.cos((2 * np.arange(1, nq + 1) - 1) * np.pi / 2 / nq)
    return x, np.pi / nq


def cutoff_fn(name):
    """SO3.py:378-464 (value, derivative)."""
    def cosine(R, rc, d=False):
        return (-0.5 * np.pi / rc * np.sin(np.pi * R / rc)) if d else 0.5 * (np.cos(np.pi * R / rc) + 1.0)

    def tanh(R, rc, d=False):
        if d:
            t2 = np.tanh(1 - R / rc) ** 2
            return -(3 / rc) * t2 * (1 - t2)
        return np.tanh(1 - R / rc) ** 3

    def poly1(R, rc, d=False):
        if d:
            return (6 / rc ** 2) * R * (R / rc - 1)
        x = R / rc
        return x ** 2 * (2 * x - 3) + 1

    def poly2(R, rc, d=False):
        x = R / rc
        return (-30 / rc) * (x ** 2 * (x - 1) ** 2) if d else x ** 3 * (x * (15 - 6 * x) - 10) + 1

    def poly3(R, rc, d=False):
        x = R / rc
        return (140 / rc) * (x ** 3 * (x - 1) ** 3) if d else x ** 4 * (x * (x * (20 * x - 70) + 84) - 35) + 1

    def poly4(R, rc, d=False):
        x = R / rc
        return (-630 / rc) * (x ** 4 * (x - 1) ** 4) if d else x ** 5 * (x * (x * (x * (315 - 70 * x) - 540) + 420)
                                                                        - 126) + 1

    def expo(R, rc, d=False):
        x = R / rc
        e = np.exp(1 - 1 / (1 - x ** 2))
        return 2 * x * e / (1 + x ** 2) ** 2 if d else e

    def unity(R, rc, d=False):
        return np.ones(len(R))

    return {"cosine": cosine, "tanh": tanh, "poly1": poly1, "poly2": poly2, "poly3": poly3, "poly4": poly4,
            "exp": expo, "unity": unity}[name]


def compute_dcs(pos, nmax, lmax, rcut, alpha, cutoff, derivative=True):
    """C [Nnb, nmax, lmax+1, 2lmax+1] and dC [..., 3] (SO3.py:504-694)."""
    w = W(nmax)
    Ris = np.linalg.norm(pos, axis=1)
    upos = pos / Ris[:, None]
    GCQ, weight = gauss_chebyshev(nmax, lmax)
    weight *= rcut / 2
    Q = rcut / 2 * (GCQ + 1)
    args = 2 * alpha * np.outer(Ris, Q)
    Gs = np.zeros((nmax, len(Q)))
    for n in range(1, nmax + 1):
        Gs[n - 1, :] = g(Q, n, nmax, rcut, w) * weight
    Bes = np.zeros((len(Ris), len(Q), lmax + 1))
    dBes = np.zeros_like(Bes)
    for l in range(lmax + 1):
        Bes[:, :, l] = spherical_in(l, args)
        if derivative:
            dBes[:, :, l] = spherical_in(l, args, derivative=True)
    Gs *= Q ** 2 * np.exp(-alpha * Q ** 2) * np.sqrt(1 - GCQ ** 2)
    integral = np.einsum('ij,kjl->kil', Gs, Bes)
    exparray = 4 * np.pi * np.exp(-alpha * Ris ** 2)
    cut = cutoff(Ris, rcut)
    thetas = np.arccos(pos[:, 2] / Ris)
    phis = np.arctan2(pos[:, 1], pos[:, 0])
    if not derivative:
        msize = 2 * lmax + 1
        ylms = np.zeros((len(Ris), lmax + 1, msize), dtype=np.complex128)
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                ylms[:, l, msize // 2 + m] = sph_harm(m, l, phis, thetas)
        C = np.einsum('ijk,ilj->iljk', ylms, integral)
        C = np.einsum('i,ijkl->ijkl', exparray * cut, C)
        return C, None
    gradB = np.einsum('ijk,in->ijkn', dBes, upos) * 2 * alpha
    gradB = np.einsum('ijkn,j->ijkn', gradB, Q)
    gint = np.einsum('ij,kjlm->kilm', Gs, gradB)
    gradexp = (-2 * alpha * Ris * exparray)[:, None] * upos
    gcut = np.einsum('i,in->in', cutoff(Ris, rcut, True), upos)
    Msize = 2 * (lmax + 1) + 1
    msize = 2 * lmax + 1
    ylms = np.zeros((len(Ris), lmax + 2, Msize), dtype=np.complex128)
    gy = np.zeros((len(Ris), lmax + 1, msize, 3), dtype=np.complex128)
    for l in range(lmax + 2):
        for m in range(-l, l + 1):
            ylms[:, l, Msize // 2 + m] = sph_harm(m, l, phis, thetas)
    for l in range(1, lmax + 1):
        for m in range(-l, l + 1):
            mi = msize // 2 + m
            Mi = Msize // 2 + m
            x0 = -np.sqrt(((l + 1) ** 2 - m ** 2) / (2 * l + 1) / (2 * l + 3)) * l * ylms[:, l + 1, Mi] / Ris
            if abs(m) <= l - 1:
                x0 += np.sqrt((l ** 2 - m ** 2) / (2 * l - 1) / (2 * l + 1)) * (l + 1) * ylms[:, l - 1, Mi] / Ris
            xp = -np.sqrt((l + m + 1) * (l + m + 2) / 2 / (2 * l + 1) / (2 * l + 3)) * l * ylms[:, l + 1, Mi + 1] / Ris
            if abs(m + 1) <= l - 1:
                xp -= np.sqrt((l - m - 1) * (l - m) / 2 / (2 * l - 1) / (2 * l + 1)) * (l + 1) * ylms[:, l - 1,
                                                                                                    Mi + 1] / Ris
            xm = -np.sqrt((l - m + 1) * (l - m + 2) / 2 / (2 * l + 1) / (2 * l + 3)) * l * ylms[:, l + 1, Mi - 1] / Ris
            if abs(m - 1) <= l - 1:
                xm -= np.sqrt((l + m - 1) * (l + m) / 2 / (2 * l - 1) / (2 * l + 1)) * (l + 1) * ylms[:, l - 1,
                                                                                                    Mi - 1] / Ris
            gy[:, l, mi, 0] = 1 / np.sqrt(2) * (xm - xp)
            gy[:, l, mi, 1] = 1j / np.sqrt(2) * (xm + xp)
            gy[:, l, mi, 2] = x0
    ylms = ylms[:, 0:lmax + 1, 1:1 + 2 * lmax + 1]
    YI = np.einsum('ijk,ilj->iljk', ylms, integral)
    dYI = np.einsum('ijkn,ilj->iljkn', gy, integral)
    YdI = np.einsum('ijk,iljn->iljkn', ylms, gint)
    C = np.einsum('i,ijkl->ijkl', exparray, YI)
    dC = np.einsum('in,ijkl->ijkln', gradexp, YI) + np.einsum('ijkln,i->ijkln', dYI + YdI, exparray)
    dC *= cut[:, None, None, None, None]
    dC += np.einsum('in,ijkl->ijkln', gcut, C)
    C *= cut[:, None, None, None]
    return C, dC


def calculate(positions, cell, pbc, numbers, nmax=3, lmax=3, rcut=3.5, alpha=2.0, derivative=True, stress=False,
              cutoff="cosine", weight_on=False, chunk=4096, centres=None):
    """SO3.calculate (SO3.py:187-291) -> dict x, dxdr, rdxdr, elements, seq.

    centres (derivative=False only): evaluate only the images of these
    centre atoms and return x of these centres, in the given order (the
    per-centre sums are independent, so the values are those of the full
    calculation)."""
    positions = np.asarray(positions, dtype=np.float64)
    numbers = np.asarray(numbers)
    N = len(positions)
    ii, jj, TT, rr = _oracle.nlist(positions, cell, pbc, rcut)
    if centres is not None:
        if derivative:
            raise ValueError("centres is for derivative=False")
        centres = np.asarray(centres, dtype=np.int64)
        keep = np.isin(ii, centres)
        ii, jj, rr = ii[keep], jj[keep], rr[keep]
    wts = numbers[jj].astype(np.int64)
    if weight_on:
        wts = np.where(numbers[jj] != numbers[ii], -wts, wts)
    # seq rows (SO3.py:358-369)
    seq = []
    for i in range(N):
        js = sorted(set(jj[ii == i].tolist()) | {i})
        seq += [(i, j) for j in js]
    seq = np.array(seq, dtype=np.int64).reshape(-1, 2)
    row_of = {(int(a), int(b)): k for k, (a, b) in enumerate(seq)}
    ls = np.arange(lmax + 1)
    norm = np.sqrt(2 * np.sqrt(2) * np.pi / np.sqrt(2 * ls + 1))
    tri = np.tril_indices(nmax, k=0)
    ncoefs = nmax * (nmax + 1) // 2 * (lmax + 1)
    cut = cutoff_fn(cutoff)
    ctot = np.zeros((N, nmax, lmax + 1, 2 * lmax + 1), dtype=np.complex128)
    for s in range(0, len(ii), chunk):
        e = slice(s, s + chunk)
        C, _ = compute_dcs(rr[e], nmax, lmax, rcut, alpha, cut, derivative=False)
        C = C * wts[e][:, None, None, None]
        C = np.einsum('inlm,l->inlm', C, norm)
        np.add.at(ctot, ii[e], C)
    P = np.einsum('aijk,aljk->ailj', ctot, np.conj(ctot)).real
    x = P[:, tri[0], tri[1], :].reshape(N, ncoefs)
    if centres is not None:
        return {"x": x[centres], "dxdr": None, "rdxdr": None, "elements": None, "seq": None}
    if not derivative:
        return {"x": x, "dxdr": None, "rdxdr": None, "elements": None, "seq": seq}
    dxdr = np.zeros((len(seq), ncoefs, 3))
    pst = np.zeros((len(seq), ncoefs, 3, 3)) if stress else None
    rows = np.array([row_of[(int(a), int(b))] for a, b in zip(ii, jj)], dtype=np.int64)
    nsite = np.array([row_of[(i, i)] for i in range(N)], dtype=np.int64)
    for s in range(0, len(ii), chunk):
        e = slice(s, s + chunk)
        _, dC = compute_dcs(rr[e], nmax, lmax, rcut, alpha, cut, derivative=True)
        dC = dC * wts[e][:, None, None, None, None]
        dC = np.einsum('inlmj,l->inlmj', dC, norm)
        ct = ctot[ii[e]]
        dP = np.einsum('wijkn,wljk->wiljn', dC, np.conj(ct))
        dP = dP + np.conj(np.transpose(dP, axes=[0, 2, 1, 3, 4]))
        dP = dP.real[:, tri[0], tri[1], :, :].reshape(-1, ncoefs, 3)
        np.add.at(dxdr, rows[e], dP)
        if stress:
            Ri = positions[ii[e]]
            Rj = rr[e] + Ri
            np.add.at(pst, rows[e], -np.einsum('wa,wkb->wkab', Rj, dP))
            np.add.at(pst, nsite[ii[e]], np.einsum('wa,wkb->wkab', Ri, dP))
    # (i,i) row: minus the sum of all rows of the centre (SO3.py:262-265)
    tot = np.zeros((N, ncoefs, 3))
    np.add.at(tot, seq[:, 0], dxdr)
    dxdr[nsite] -= tot
    out = {"x": x, "dxdr": dxdr, "seq": seq}
    if stress:
        vol = abs(np.linalg.det(np.asarray(cell)))
        out["rdxdr"] = -pst / vol
    else:
        out["rdxdr"] = None
    return out
