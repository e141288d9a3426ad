"""CPU oracle for the descriptor hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
package (``pyxtal_ff_amd``) never imports it.

Contents
  * ``so4_oracle.c``  -- statement-level C restatement of the SNAP/SO4 numba
    kernels (``pyxtal_ff/descriptors/SNAP.py:336-1762``).
  * ``nlist_oracle.c`` -- brute-force image enumeration restating ASE's
    ``NeighborList`` semantics as used at ``SNAP.py:250-252``.
  * ``so3_oracle.py`` -- NumPy/SciPy restatement of ``SO3.py``.
  * this module -- the reference's host-side bookkeeping (``seq``, padded
    ``neighborlist``, weights; ``SNAP.py:240-317``) and the per-atom rdxdr
    reduction (``pyxtal_ff/utilities/__init__.py:226-235``).

Pinning (see tests/test_oracle_*.py): Clebsch-Gordan values against exact
Racah sums in rational arithmetic, U arrays against the closed-form Wigner-D
polynomial, x against the reference-produced descriptor ranges shipped in the
reference's examples (NN_weights.txt Scale0/Scale1), dx/dr and rdxdr against
central finite differences, rotation/translation/permutation invariances and
the SO4 <-> SNAP Z^3 relation.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
_LIB_PATH = os.path.join(_BUILD, "liboracle.so")
_lib = None


def build(force=False):
    """Compile the C restatements into oracle/_build/liboracle.so."""
    os.makedirs(_BUILD, exist_ok=True)
    srcs = [os.path.join(_HERE, "so4_oracle.c"), os.path.join(_HERE, "nlist_oracle.c")]
    if not force and os.path.exists(_LIB_PATH):
        mt = os.path.getmtime(_LIB_PATH)
        if all(os.path.getmtime(s) <= mt for s in srcs):
            return _LIB_PATH
    objs = []
    for s, flags in ((srcs[0], ["-O3", "-march=x86-64-v3", "-fopenmp"]), (srcs[1], ["-O2", "-ffp-contract=off"])):
        o = os.path.join(_BUILD, os.path.basename(s)[:-2] + ".o")
        subprocess.check_call(["gcc", "-std=gnu99", "-fPIC", "-c", s, "-o", o] + flags)
        objs.append(o)
    tmp = _LIB_PATH + ".tmp%d" % os.getpid()
    subprocess.check_call(["gcc", "-shared", "-fopenmp", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, _LIB_PATH)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.so4_oracle_compute.argtypes = [P, i64, P, i64, i64, P, P, P, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, P, P, P, P]
        L.so4_oracle_compute.restype = ctypes.c_int
        L.so4_oracle_compute_mt.argtypes = L.so4_oracle_compute.argtypes + [ctypes.c_int]
        L.so4_oracle_compute_mt.restype = ctypes.c_int
        L.so4_oracle_sizes.argtypes = [ctypes.c_int, P, P, P, P]
        L.so4_oracle_tables.argtypes = [ctypes.c_int, P, P, P, P]
        L.so4_oracle_uarray.argtypes = [ctypes.c_int] + [ctypes.c_double] * 5 + [P, P]
        L.so4_oracle_cutoff.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, P, P]
        L.so4_oracle_cutoff.restype = None
        L.nlist_oracle.argtypes = [P, i64, P, P, ctypes.c_double, P, P, P, P]
        L.nlist_oracle.restype = i64
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ----------------------------------------------------------------- tables
def so4_sizes(twol):
    v = [ctypes.c_int() for _ in range(4)]
    lib().so4_oracle_sizes(int(twol), *[ctypes.byref(x) for x in v])
    return dict(ncg=v[0].value, nu=v[1].value, nz=v[2].value, nb=v[3].value)


def so4_tables(twol):
    s = so4_sizes(twol)
    cg = np.zeros(s["ncg"])
    idxz = np.zeros((s["nz"], 10), dtype=np.int32)
    idxb = np.zeros((s["nb"], 3), dtype=np.int32)
    idxu = np.zeros(twol + 1, dtype=np.int32)
    lib().so4_oracle_tables(int(twol), _ptr(cg), _ptr(idxz), _ptr(idxb), _ptr(idxu))
    return dict(cglist=cg, idxz=idxz, idxb=idxb, idxu_block=idxu, **s)


def uarray(twol, x, y, z, psi, r, derivative=False):
    nu = so4_sizes(twol)["nu"]
    u = np.zeros(2 * nu)
    du = np.zeros(6 * nu) if derivative else None
    lib().so4_oracle_uarray(int(twol), x, y, z, psi, r, _ptr(u), _ptr(du))
    u = u[0::2] + 1j * u[1::2]
    if derivative:
        du = (du[0::2] + 1j * du[1::2]).reshape(nu, 3)
        return u, du
    return u


def cutoff(name, r, rc):
    """(fc, dfc) of the SNAP cutoff ``name`` at separation r (SNAP.py:512-588)."""
    f = ctypes.c_double()
    df = ctypes.c_double()
    lib().so4_oracle_cutoff(CUTOFF_IDS[name], float(r), float(rc), ctypes.byref(f), ctypes.byref(df))
    return f.value, df.value


# ---------------------------------------------------------- neighbour list
def nlist(positions, cell, pbc, rc):
    """All images (i, j, T, r) with |r| < rc, ordered by (i, j, T)."""
    pos = np.ascontiguousarray(positions, dtype=np.float64)
    cel = np.ascontiguousarray(cell, dtype=np.float64)
    pb = np.ascontiguousarray(np.asarray(pbc, dtype=bool).astype(np.int32).reshape(3))
    n = len(pos)
    L = lib()
    cnt = L.nlist_oracle(_ptr(pos), n, _ptr(cel), _ptr(pb), float(rc), None, None, None, None)
    ii = np.zeros(cnt, dtype=np.int64)
    jj = np.zeros(cnt, dtype=np.int64)
    TT = np.zeros((cnt, 3), dtype=np.int32)
    rr = np.zeros((cnt, 3))
    L.nlist_oracle(_ptr(pos), n, _ptr(cel), _ptr(pb), float(rc), _ptr(ii), _ptr(jj), _ptr(TT), _ptr(rr))
    return ii, jj, TT, rr


def reference_layout(positions, cell, pbc, rc, atom_weights, site_weights):
    """Restates SNAP.py:240-317: seq rows (i ascending, unique j ascending,
    (i,i) always present), padded neighborlist [P, M, 3], pair weights [P, M]
    (image order within a row = (T0,T1,T2) ascending), site weights."""
    n = len(positions)
    ii, jj, TT, rr = nlist(positions, cell, pbc, rc)
    seq = []
    groups = []
    start = 0
    for i in range(n):
        end = start
        while end < len(ii) and ii[end] == i:
            end += 1
        js = jj[start:end]
        uniq = sorted(set(js.tolist()) | {i})
        for j in uniq:
            sel = np.nonzero(js == j)[0] + start
            seq.append((i, j))
            groups.append(sel)
        start = end
    P = len(seq)
    M = max([len(g) for g in groups] + [0])
    nl = np.zeros((P, max(M, 1) if P else 0, 3)) if M > 0 else np.zeros((P, 0, 3))
    w = np.zeros((P, M))
    for p, g in enumerate(groups):
        nl[p, :len(g)] = rr[g]
        w[p, :len(g)] = np.asarray(atom_weights, dtype=np.float64)[jj[g]] if len(g) else 0
    return dict(center_atoms=np.ascontiguousarray(positions, dtype=np.float64),
                neighborlist=np.ascontiguousarray(nl[:, :M]), seq=np.asarray(seq, dtype=np.int64).reshape(-1, 2),
                neighbor_w=np.ascontiguousarray(w), site_w=np.asarray(site_weights, dtype=np.float64),
                images=(ii, jj, TT, rr))


# ------------------------------------------------------------------- SO4
CUTOFF_IDS = {"cosine": 1, "tanh": 2, "poly1": 3, "poly2": 4, "poly3": 5, "poly4": 6, "exp": 7}


def so4_components(layout, twol, rcut, norm=False, derivative=True, stress=True, cutoff="cosine",
                   rfac0=0.99363, want_z=False, threads=1):
    """get_bispectrum_components (SNAP.py:1289-1762) on a reference layout;
    threads > 1 processes centres on that many OpenMP threads (same values)."""
    ca = layout["center_atoms"]
    nl = layout["neighborlist"]
    seq = layout["seq"]
    nw = layout["neighbor_w"]
    sw = layout["site_w"]
    N = len(ca)
    P, M = nl.shape[0], nl.shape[1]
    s = so4_sizes(twol)
    nb, nz = s["nb"], s["nz"]
    bl = np.zeros((N, nb))
    db = np.empty((P, nb, 3)) if derivative else None  # zeroed (in parallel) by the C code
    bs = np.empty((P, nb, 3, 3)) if (derivative and stress) else None
    zl = np.zeros((N, nz, 2)) if want_z else None
    nlc = np.ascontiguousarray(nl if M > 0 else np.zeros((P, 1, 3)))
    nwc = np.ascontiguousarray(nw if M > 0 else np.zeros((P, 1)))
    rc = lib().so4_oracle_compute_mt(_ptr(np.ascontiguousarray(ca)), N, _ptr(nlc), P, max(M, 1),
                                  _ptr(np.ascontiguousarray(seq)), _ptr(nwc),
                                  _ptr(np.ascontiguousarray(sw)), int(twol), float(rcut), int(bool(norm)),
                                  int(bool(derivative)), int(bool(stress)), CUTOFF_IDS[cutoff], float(rfac0),
                                  _ptr(bl), _ptr(db), _ptr(bs), _ptr(zl), int(threads))
    if rc != 0:
        raise ValueError("not implemented")
    out = dict(blist=bl, dblist=db, bstress=bs)
    if want_z:
        out["zlist"] = zl[..., 0] + 1j * zl[..., 1]
    return out


def snap_calculate(positions, cell, pbc, symbols, weights, lmax=3, rcut=3.5, derivative=True, stress=False,
                   normalize_U=False, cutoff_function="cosine", rfac0=0.99363, threads=1):
    """SNAP.py::SO4_Bispectrum.calculate (SNAP.py:200-238): dict x, dxdr,
    rdxdr (per pair row, -bstress/V), elements, seq."""
    twol = int(round(2 * lmax))
    w_atom = np.array([weights[s] for s in symbols], dtype=np.float64)
    lay = reference_layout(positions, cell, pbc, rcut, w_atom, np.ones(len(positions)))
    out = so4_components(lay, twol, rcut, normalize_U, derivative, stress, cutoff_function, rfac0, threads=threads)
    vol = abs(np.linalg.det(np.asarray(cell, dtype=np.float64)))
    return _assemble(out, lay, symbols, derivative, stress, vol)


def so4_calculate(positions, cell, pbc, numbers, symbols, lmax=3, rcut=3.5, derivative=True, stress=False,
                  normalize_U=False, cutoff_function="cosine"):
    """SO4.py::SO4_Bispectrum.calculate (SO4.py:171-209): weights Z_j, self
    weight Z_i, psi = pi r / rc."""
    twol = 2 * int(lmax)
    z = np.asarray(numbers, dtype=np.float64)
    lay = reference_layout(positions, cell, pbc, rcut, z, z)
    out = so4_components(lay, twol, rcut, normalize_U, derivative, stress, cutoff_function, 1.0)
    vol = abs(np.linalg.det(np.asarray(cell, dtype=np.float64)))
    return _assemble(out, lay, symbols, derivative, stress, vol)


def _assemble(out, lay, symbols, derivative, stress, vol):
    if derivative:
        d = {"x": out["blist"], "dxdr": out["dblist"], "elements": list(symbols), "seq": lay["seq"]}
        d["rdxdr"] = -out["bstress"] / vol if stress else None
    else:
        d = {"x": out["blist"], "dxdr": None, "elements": list(symbols), "rdxdr": None}
    return d


def rdxdr_per_atom(d):
    """utilities/__init__.py:226-235: per-atom sum of the per-row rdxdr,
    reshaped to 9 and reduced to Voigt [xx, yy, zz, xy, xz, yz]."""
    N, L = d["x"].shape
    out = np.zeros((N, L, 3, 3))
    np.add.at(out, d["seq"][:, 0], d["rdxdr"])
    return out.reshape(N, L, 9)[:, :, [0, 4, 8, 1, 2, 5]]


def dense_dxdr(d):
    """[N, N, L, 3] dense layout the reference tests build (test_all.py:222-227)."""
    N, L = d["x"].shape
    arr = np.zeros((N, N, L, 3))
    np.add.at(arr, (d["seq"][:, 0], d["seq"][:, 1]), d["dxdr"])
    return arr
