"""Seeded synthetic structures for the benchmark and parity workloads.

  * ``diamond_si(reps, sigma, seed)`` -- diamond-cubic Si (a = 5.43 A),
    reps^3 conventional cells, N(0, sigma) displacements (64 atoms at
    reps = 2, 4096 atoms at reps = 8);
  * ``random_sio_batch(n, seed)`` -- periodic Si/O cells, N uniform in
    [8, 256], O fraction uniform in [0, 0.67], density 0.066 / A^3, +-5 %
    random strain, 1.5 A minimum-distance rejection, atoms sorted by element;
  * ``amorphous_sio2(n_si, n_o, seed)`` -- random SiO2 at 2.2 g/cm^3 with a
    1.6 A minimum distance;
  * ``nacl_cluster(angle)`` -- the 8-atom rocksalt NaCl cluster in a 17.22 A
    periodic box used by the reference's unit tests (test_all.py:23-42).
"""
import numpy as np

from .structure import Structure

DIAMOND_BASIS = np.array([[0, 0, 0], [0.25, 0.25, 0.25], [0, 0.5, 0.5], [0.25, 0.75, 0.75],
                          [0.5, 0, 0.5], [0.75, 0.25, 0.75], [0.5, 0.5, 0], [0.75, 0.75, 0.25]])


def diamond_si(reps=2, sigma=0.05, seed=0, a=5.43):
    cells = []
    for i in range(reps):
        for j in range(reps):
            for k in range(reps):
                cells.append(DIAMOND_BASIS + np.array([i, j, k]))
    frac = np.concatenate(cells, axis=0)
    pos = frac * a
    rng = np.random.default_rng(seed)
    pos = pos + rng.normal(0.0, sigma, pos.shape)
    L = reps * a
    # keep atoms inside the cell (the image set does not depend on it)
    pos = np.mod(pos, L)
    return Structure(pos, np.eye(3) * L, numbers=[14] * len(pos), pbc=True)


def _random_positions(rng, cell, n, dmin, max_tries=20000):
    """Uniform positions in the cell, rejecting any within dmin (minimum
    image) of an accepted one."""
    inv = np.linalg.inv(cell)
    pts = np.zeros((n, 3))
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > max_tries * max(n, 1):
            raise RuntimeError("could not place atoms at the requested minimum distance")
        p = rng.random(3) @ cell
        if k:
            d = p - pts[:k]
            d = d - np.round(d @ inv) @ cell
            if np.min(np.einsum("ij,ij->i", d, d)) < dmin * dmin:
                continue
        pts[k] = p
        k += 1
    return pts


def random_sio_structure(rng, nmin=8, nmax=256, density=0.066, dmin=1.5):
    n = int(rng.integers(nmin, nmax + 1))
    fo = rng.uniform(0.0, 0.67)
    n_o = int(round(fo * n))
    n_si = n - n_o
    L = (n / density) ** (1.0 / 3.0)
    strain = np.eye(3) + rng.uniform(-0.05, 0.05, (3, 3))
    cell = np.eye(3) * L @ strain
    pos = _random_positions(rng, cell, n, dmin)
    numbers = [14] * n_si + [8] * n_o  # element-sorted
    return Structure(pos, cell, numbers=numbers, pbc=True)


def random_sio_batch(nstruct=1024, seed=0, **kw):
    rng = np.random.default_rng(seed)
    return [random_sio_structure(rng, **kw) for _ in range(nstruct)]


def amorphous_sio2(n_si=4000, n_o=8000, seed=0, density_gcc=2.2, dmin=1.6):
    """Random SiO2 at the given mass density with a minimum distance (cubic,
    periodic); a cell grid accelerates the rejection test."""
    rng = np.random.default_rng(seed)
    mass = (n_si * 28.0855 + n_o * 15.999) / 6.02214076e23  # g
    L = (mass / density_gcc * 1e24) ** (1.0 / 3.0)  # A
    n = n_si + n_o
    nb = max(1, int(L // dmin))
    h = L / nb
    cells = {}
    pts = np.zeros((n, 3))
    k = 0
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    while k < n:
        cand = rng.random((256, 3)) * L
        for p in cand:
            c = tuple(int(v) % nb for v in p // h)
            near = []
            for o in offs:
                near.extend(cells.get(((c[0] + o[0]) % nb, (c[1] + o[1]) % nb, (c[2] + o[2]) % nb), ()))
            if near:
                d = p - pts[near]
                d -= np.round(d / L) * L
                if np.min(np.einsum("ij,ij->i", d, d)) < dmin * dmin:
                    continue
            cells.setdefault(c, []).append(k)
            pts[k] = p
            k += 1
            if k == n:
                break
    numbers = [14] * n_si + [8] * n_o
    return Structure(pts, np.eye(3) * L, numbers=numbers, pbc=True)


def nacl_cluster(angle=1.0, axis="x"):
    """Rocksalt NaCl (a = 5.691694) conventional cell as a cluster, rotated by
    `angle` degrees about x, boxed in a periodic 17.22 A cube (the reference's
    unit-test fixture, test_all.py:23-42)."""
    a = 5.691694
    # ase.build.bulk('NaCl', 'rocksalt', cubic=True) then sort(tags=[0,4,1,5,2,6,3,7])
    na = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]]) * a
    cl = np.array([[0.5, 0, 0], [0.5, 0.5, 0.5], [0, 0, 0.5], [0, 0.5, 0]]) * a
    pos = np.concatenate([na, cl])
    numbers = [11] * 4 + [17] * 4
    th = np.radians(angle)
    c, s = np.cos(th), np.sin(th)
    if axis == "x":
        R = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    elif axis == "y":
        R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    else:
        R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    pos = pos @ R.T
    return Structure(pos, np.eye(3) * 17.22, numbers=numbers, pbc=True)


def rotate(struct, angle, axis="x"):
    th = np.radians(angle)
    c, s = np.cos(th), np.sin(th)
    R = {"x": np.array([[1, 0, 0], [0, c, -s], [0, s, c]]),
         "y": np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]]),
         "z": np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])}[axis]
    out = struct.copy()
    out.positions = struct.positions @ R.T
    return out
