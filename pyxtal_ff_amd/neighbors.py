"""Device neighbour list (the reference's ``build_neighbor_list``).

Replaces ``pyxtal_ff/descriptors/SNAP.py:240-317`` (ASE ``NeighborList`` with
cutoffs ``[rc/2]*N``, ``self_interaction=False``, ``bothways=True``,
``skin=0``, then the per-centre regrouping into ``seq`` rows) for a batch of
structures in three device phases (count, fill, rows).  Everything stays in
HBM; only the two totals (images, rows) come back to the host to size the
outputs.

Per image e: centre ``img_atom``, neighbour ``img_j``, translation ``img_T``,
separation ``img_r`` (= pos_j + T.cell - pos_i), its seq row ``img_row`` and
the reverse image ``img_rev`` (j, i, -T).  Per row: ``seq`` (i, j),
``row_img`` (image range) and ``row_rev``.  Per atom: ``img_start``,
``row_start`` and ``nsite`` (the (i, i) row).
"""
import math

import numpy as np
import torch

from . import _native as nat


def binning_lattice(cell, pbc):
    """The lattice the atoms are binned in: the cell with the rows of its
    non-periodic directions replaced by unit vectors orthogonal to the
    periodic rows.  Along a non-periodic direction there is one bin and no
    translation, so those rows only have to make the matrix invertible --
    a cluster may come with a zero or degenerate cell there, which ASE's
    NeighborList accepts (SNAP.py:250-252, SO3.py:326-332).  Separations
    are still formed with the structure's own cell."""
    c = np.array(cell, dtype=np.float64).reshape(3, 3)
    per = [d for d in range(3) if pbc[d]]
    if len(per) == 3:
        return c
    non = [d for d in range(3) if not pbc[d]]
    if per:
        basis = c[per]
        if np.linalg.matrix_rank(basis) < len(per):
            raise ValueError("the cell vectors of the periodic directions are linearly dependent")
        comp = np.linalg.svd(basis)[2][len(per):]  # orthonormal complement of the periodic rows
    else:
        comp = np.eye(3)
    for k, d in enumerate(non):
        c[d] = comp[k]
    return c


def _geometry(cell, pbc, rc):
    """Bin grid of one structure: nb_d = floor(h_d / rc) (>= 1) bins along
    lattice direction d, stencil half-width S_d = ceil(rc nb_d / h_d), where
    h_d is the spacing of the lattice planes normal to direction d.  Returns
    the inverse of the binning lattice (``binning_lattice``) and the grid."""
    inv = np.linalg.inv(binning_lattice(cell, pbc))
    geom = np.zeros(8, dtype=np.int32)
    for d in range(3):
        if not pbc[d]:
            geom[d], geom[3 + d] = 1, 0
            continue
        h = 1.0 / np.linalg.norm(inv[:, d])
        nb = max(1, int(math.floor(h / rc)))
        nb = min(nb, 1024)
        S = int(math.ceil(rc * nb / h * (1 + 1e-12) + 1e-12))
        geom[d], geom[3 + d] = nb, S
    return inv, geom


class NeighborListBuilder:
    """Structure metadata (cells, bin grids, atom ranges) and workspace for
    repeated neighbour-list builds of the same set of structures -- e.g. the
    steps of an MD run or a training epoch.  ``build(pos)`` redoes the whole
    list from the given positions; only the metadata is reused."""

    # first image capacity of a builder: max(MIN_CAPACITY, CAPACITY_PER_ATOM x atoms)
    MIN_CAPACITY = 1024
    CAPACITY_PER_ATOM = 48
    # images a centre keeps from the count for the fill (more: enumerated again)
    SLOT = 64

    def __init__(self, natoms_per_struct, cells, pbcs, rc, device="cuda"):
        counts = np.asarray(natoms_per_struct, dtype=np.int64).reshape(-1)
        cells = np.asarray(cells, dtype=np.float64).reshape(-1, 3, 3)
        pbcs = np.asarray(pbcs, dtype=bool).reshape(-1, 3)
        nst = len(counts)
        if cells.shape[0] != nst or pbcs.shape[0] != nst:
            raise ValueError("positions, cells and pbcs must describe the same structures")
        self.rc = float(rc)
        atom_start = np.zeros(nst + 1, dtype=np.int64)
        atom_start[1:] = np.cumsum(counts)
        self.natoms = int(atom_start[-1])
        self.nstruct = nst
        self.atom_start_host = atom_start
        invs = np.zeros((nst, 9))
        geoms = np.zeros((nst, 8), dtype=np.int32)
        off = 0
        for s in range(nst):
            inv, g = _geometry(cells[s], pbcs[s], self.rc)
            invs[s] = inv.reshape(9)
            g[6] = off
            off += int(g[0]) * int(g[1]) * int(g[2])
            geoms[s] = g
        self.nbins = off
        dev = torch.device(device)
        self.device = dev
        f64 = dict(dtype=torch.float64, device=dev)
        i64 = dict(dtype=torch.int64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.atom_struct = torch.as_tensor(np.repeat(np.arange(nst), counts), **i64)
        self.atom_start = torch.as_tensor(atom_start, **i64)
        self.cells = torch.as_tensor(cells.reshape(nst, 9), **f64).contiguous()
        self.inv_cells = torch.as_tensor(invs, **f64).contiguous()
        self.pbc = torch.as_tensor(pbcs.astype(np.int32), **i32).contiguous()
        self.geom = torch.as_tensor(geoms, **i32).contiguous()
        self.volumes = np.abs(np.linalg.det(cells)) if nst else np.zeros(0)
        # a cluster's zero cell has no volume: its stress rows come out as the
        # reference's -bstress / V does (inf / nan), everything else is finite
        with np.errstate(divide="ignore"):
            self.inv_vol = torch.as_tensor(np.repeat(1.0 / self.volumes, counts), **f64)
        n = self.natoms
        nb = max(self.nbins, 1)
        sl = max(n, 1) * self.SLOT
        self._ws = dict(bin_count=torch.empty(nb, **i32), bin_start=torch.empty(nb + 1, **i32),
                        atom_bin=torch.empty(max(n, 1), **i32), wrap=torch.empty(max(n, 1) * 3, **i32),
                        bin_atoms=torch.empty(max(n, 1), **i32), slot_j=torch.empty(sl, **i64),
                        slot_T=torch.empty(sl * 3, **i32), slot_r=torch.empty(sl * 3, dtype=torch.float64, device=dev))

        # image capacity of the next build: the arrays are sized before the
        # image count is known on the host (no synchronisation in build);
        # it grows when a build overflows it
        self.img_capacity = max(self.MIN_CAPACITY, self.CAPACITY_PER_ATOM * n)

    def build(self, pos, host_count=True):
        """A DeviceNeighborList of float64 device positions pos [natoms, 3].
        host_count=False: no copy of the image count to the host (the list
        can be built inside a HIP graph capture; ``nimg`` is then not
        available and consumers size their launches by ``img_capacity``)."""
        return DeviceNeighborList(None, None, None, self.rc, pos_device=pos, builder=self, host_count=host_count)


class DeviceNeighborList:
    """Neighbour images and seq rows of a batch of structures on the GPU.

    positions: list of [N_s, 3] arrays, cells: [nstruct, 3, 3],
    pbcs: [nstruct, 3].  With ``pos_device`` (a float64 [N, 3] device tensor)
    positions are taken from HBM; ``builder`` supplies cached metadata."""

    def __init__(self, positions, cells, pbcs, rc, device="cuda", pos_device=None, builder=None, host_count=True):
        if builder is None:
            if pos_device is not None:
                if positions is not None:
                    raise ValueError("give positions or pos_device, not both")
                counts = [pos_device.shape[0]]
            else:
                positions = [np.asarray(p, dtype=np.float64).reshape(-1, 3) for p in positions]
                counts = [len(p) for p in positions]
            builder = NeighborListBuilder(counts, cells, pbcs, rc, device)
        b = builder
        self.builder = b
        self.rc = b.rc
        self.natoms, self.nstruct = b.natoms, b.nstruct
        self.atom_start_host, self.atom_struct, self.atom_start = b.atom_start_host, b.atom_struct, b.atom_start
        self.cells, self.inv_cells, self.pbc, self.geom = b.cells, b.inv_cells, b.pbc, b.geom
        self.volumes, self.inv_vol = b.volumes, b.inv_vol
        f64 = dict(dtype=torch.float64, device=b.device)
        i64 = dict(dtype=torch.int64, device=b.device)
        i32 = dict(dtype=torch.int32, device=b.device)
        if pos_device is not None:
            if pos_device.dtype != torch.float64 or not pos_device.is_cuda or pos_device.shape != (b.natoms, 3):
                raise ValueError("pos_device must be a float64 [natoms, 3] device tensor")
            self.pos = pos_device.contiguous()
        else:
            pos_all = np.concatenate(positions, axis=0) if b.nstruct else np.zeros((0, 3))
            self.pos = torch.as_tensor(pos_all, **f64).contiguous()
        self._host_count = bool(host_count)
        self._build(b.nbins, f64, i64, i32)

    def _build(self, nbins, f64, i64, i32):
        # per-build counts (a later build must not overwrite this list's)
        self._img_count = torch.empty(max(self.natoms, 1), **i64)
        self._count()
        self.generation = 0
        self._fill(self.builder.img_capacity)

    def _count(self):
        """Bin the atoms (into the builder's shared workspace) and count the
        images of every centre; centres with at most SLOT images keep them,
        sorted, in the builder's slot arrays for the fill that follows (every
        count is followed by its fill on the same stream)."""
        ws = self.builder._ws
        p = nat.ptr
        self._row_count = torch.empty(max(self.natoms, 1), dtype=torch.int64, device=self.pos.device)
        nat.check(nat.lib().pxff_nlist_count_slots(
            p(self.pos), self.natoms, p(self.atom_struct), p(self.atom_start), p(self.cells), p(self.inv_cells),
            p(self.pbc), p(self.geom), max(self.builder.nbins, 1), self.rc, p(ws["bin_count"]), p(ws["bin_start"]),
            p(ws["atom_bin"]), p(ws["wrap"]), p(ws["bin_atoms"]), p(self._img_count), self.builder.SLOT,
            p(ws["slot_j"]), p(ws["slot_T"]), p(ws["slot_r"]), p(self._row_count), nat.stream()),
            "pxff_nlist_count_slots")

    def _fill(self, cap):
        """Scan (clamped to cap), fill and reverse phases into arrays of cap
        images.  The unclamped total is copied to pinned host memory
        asynchronously; ``nimg`` waits for it."""
        n = self.natoms
        dev = self.pos.device
        i64 = dict(dtype=torch.int64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        f64 = dict(dtype=torch.float64, device=dev)
        s = nat.stream()
        L = nat.lib()
        p = nat.ptr
        ws = self.builder._ws
        bin_start, atom_bin, wrap, bin_atoms = ws["bin_start"], ws["atom_bin"], ws["wrap"], ws["bin_atoms"]
        self._cap = int(cap)
        m = max(self._cap, 1)
        img_start = torch.empty(n + 1, **i64)
        total = torch.empty(1, **i64)
        nat.check(L.pxff_scan_i64_capped(p(self._img_count), n, self._cap, p(img_start), p(total), s),
                  "pxff_scan_i64_capped")
        self._total = total
        if self._host_count:
            self._total_host = torch.empty(1, dtype=torch.int64, pin_memory=True)
            self._total_host.copy_(total, non_blocking=True)
            self._total_ev = torch.cuda.Event()
            self._total_ev.record()
        self._nimg = None
        self.img_start = img_start
        self.img_atom = torch.empty(m, **i64)
        self.img_j = torch.empty(m, **i64)
        self.img_T = torch.empty(m * 3, **i32)
        self.img_r = torch.empty(m * 3, **f64)
        nat.check(L.pxff_nlist_fill_slots(p(self.pos), n, p(self.atom_struct), p(self.atom_start), p(self.cells),
                                          p(self.pbc), p(self.geom), self.rc, p(bin_start), p(atom_bin), p(wrap),
                                          p(bin_atoms), p(img_start), p(self._img_count), self.builder.SLOT,
                                          p(ws["slot_j"]), p(ws["slot_T"]), p(ws["slot_r"]), p(self.img_atom),
                                          p(self.img_j), p(self.img_T), p(self.img_r), p(self._row_count), s),
                  "pxff_nlist_fill_slots")
        self.img_rev = torch.empty(m, **i64)
        # no symmetry flag here (one launch fewer per build); verify() redoes
        # this deterministic phase with the flag
        nat.check(L.pxff_nlist_rev(n, self._cap, p(img_start), p(self.img_atom), p(self.img_j), p(self.img_T),
                                   p(self.img_rev), None, s), "pxff_nlist_rev")
        self._flag = None
        self._rev_checked = False
        self._rows_built = False
        self._shift = None

    @property
    def nimg(self):
        """Number of images (waits for the count copied back by the build).
        If the build's image capacity was too small, the scan / fill /
        reverse phases are redone with a larger one first, and
        ``generation`` increases: work launched on this list before reading
        nimg used a cut list and must be redone."""
        if self._nimg is None:
            if not self._host_count:
                raise nat.PxffError("this list was built without the host image count (host_count=False)")
            self._total_ev.synchronize()
            total = int(self._total_host[0])
            if total > self._cap:
                b = self.builder
                b.img_capacity = max(b.img_capacity, total + total // 4 + 64)
                self._count()  # the shared bins may hold a later build's atoms
                self._fill(b.img_capacity)
                self.generation += 1
                self._total_ev.synchronize()
            self._nimg = total
        return self._nimg

    def resolve(self):
        """Settle the image count before anything is launched on the list:
        wait for it and, if the build overflowed its capacity, redo the
        build with a larger one (``nimg``).  Callers that read no image
        count before their first kernel call this first, so no kernel ever
        runs on a cut list.  Returns the list."""
        if self._host_count:
            self.nimg
        return self

    @property
    def img_capacity(self):
        return self._cap

    @property
    def total_device(self):
        """Unclamped image count [1] on the device (a capacity above it means
        the list is complete)."""
        return self._total

    def verify(self):
        """Raise if the reverse-image / reverse-row search failed (reads a
        device flag, i.e. synchronises)."""
        self._ensure_flag()
        if not self._rev_checked:
            p = nat.ptr
            nat.check(nat.lib().pxff_nlist_rev(self.natoms, self._cap, p(self.img_start), p(self.img_atom),
                                               p(self.img_j), p(self.img_T), p(self.img_rev), p(self._flag),
                                               nat.stream()), "pxff_nlist_rev")
            self._rev_checked = True
        fl = int(self._flag.max().item())
        if fl != 0:
            raise nat.PxffError("neighbour list is not symmetric (flag %d)" % fl)

    def _ensure_flag(self):
        if self._flag is None:
            self._flag = torch.zeros(2, dtype=torch.int32, device=self.pos.device)

    def _build_rows(self):
        if self._rows_built:
            return
        n = self.natoms
        dev = self.pos.device
        i64 = dict(dtype=torch.int64, device=dev)
        p = nat.ptr
        s = nat.stream()
        L = nat.lib()
        row_start = torch.empty(n + 1, **i64)
        nat.check(L.pxff_scan_i64(p(self._row_count), n, p(row_start), s), "pxff_scan_i64")
        nrows = int(row_start[n].item())
        r = max(nrows, 1)
        m = max(self.nimg, 1)
        self._nrows = nrows
        self._row_start = row_start
        self._seq = torch.empty(r * 2, **i64)
        self._row_img = torch.empty(r + 1, **i64)
        self._img_row = torch.empty(m, **i64)
        self._nsite = torch.empty(max(n, 1), **i64)
        self._row_rev = torch.empty(r, **i64)
        self._ensure_flag()
        flag = self._flag[1:2]
        nat.check(L.pxff_nlist_rows(n, self.nimg, nrows, p(self.img_start), p(self.img_j), p(row_start),
                                    p(self._seq), p(self._row_img), p(self._img_row), p(self._nsite),
                                    p(self._row_rev), p(flag), s), "pxff_nlist_rows")
        self._rows_built = True
        self.verify()

    # seq rows are built on first use (the adjoint path never needs them)
    @property
    def nrows(self):
        self._build_rows()
        return self._nrows

    @property
    def row_start(self):
        self._build_rows()
        return self._row_start

    @property
    def seq(self):
        self._build_rows()
        return self._seq

    @property
    def row_img(self):
        self._build_rows()
        return self._row_img

    @property
    def img_row(self):
        self._build_rows()
        return self._img_row

    @property
    def nsite(self):
        self._build_rows()
        return self._nsite

    @property
    def row_rev(self):
        self._build_rows()
        return self._row_rev

    # views in the reference's shapes
    def seq_array(self):
        return self.seq[: 2 * self.nrows].view(self.nrows, 2)

    def images(self):
        """(i, j, T, r) per image, ordered by (i, j, T)."""
        k = self.nimg
        return (self.img_atom[:k], self.img_j[:k], self.img_T[: 3 * k].view(k, 3), self.img_r[: 3 * k].view(k, 3))

    def counts(self):
        return (self.img_start[1:] - self.img_start[:-1])
