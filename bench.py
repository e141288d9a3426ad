"""Benchmark: SO4 2J=8 energy+forces throughput (atom-environments/s), the
adjoint path against the plain materialised-dxdr path on the same GPU, FP64.

Workload: diamond-cubic Si 8x8x8 supercell (4096 atoms), N(0, 0.08 A)
seeded displacements, rc = 4.9 A, SNAP 2J = 8, w_Si = 1.0, rfac0 = 0.99363,
cosine cutoff; energy from a seeded per-element MLP 55-30-30-1 (tanh, FP64).
One step = neighbour list + descriptors + network energy + forces for the
whole structure; inputs (positions) are resident in HBM when timing starts.
The same step captured once as a HIP graph and replayed
(``potential.FusedEFGraph``), the autograd variant and the plain
materialised-dxdr path are timed beside it.

Prints one JSON line (see the keys below).  ``--gpus N`` with N > 1 exits:
only one GPU's share of the work is measured.
"""
import argparse
import json
import os
import sys
import time
import traceback

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-envs/s SO4 2J=8 energy+forces: adjoint vs plain dxdr path, same GPU, FP64"
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12


# ------------------------------------------------------ algorithmic counts
def so4_counts(twol):
    """Loop-bound counts of the reference algorithm (SNAP.py:1050-1257)."""
    zterms = zpairs = 0
    nb = 0
    for l1 in range(twol + 1):
        for l2 in range(l1 + 1):
            for l in range(l1 - l2, min(twol, l1 + l2) + 1, 2):
                if l >= l1:
                    nb += 1
                mb = 0
                while 2 * mb <= l:
                    for ma in range(l + 1):
                        ma1min = max(0, (2 * ma - l - l2 + l1) // 2)
                        na = min(l1, (2 * ma - l + l2 + l1) // 2) - ma1min + 1
                        mb1min = max(0, (2 * mb - l - l2 + l1) // 2)
                        nbb = min(l1, (2 * mb - l + l2 + l1) // 2) - mb1min + 1
                        zterms += na * nbb
                        zpairs += nbb
                    mb += 1
    half = [(l // 2 + 1) * (l + 1) for l in range(twol + 1)]
    return dict(zterms=zterms, zpairs=zpairs, half=half, nhalf=sum(half), nb=nb)


def flops(twol, k_mean):
    """FP64 flops per atom-environment (conventions: complex multiply 6,
    real x complex 2, complex add 2, multiply-add 2).
      Z: 10 per inner term (cg * u1 * u2 accumulated) + 4 per outer term;
      U per neighbour: 18 per half element (two scaled complex products);
      the reference's dU per neighbour: 3 x 34 per half element, and its
      adjoint contraction 16 per half element (SNAP.py:958-1048);
      the adjoint sweep as run here (U alone, dU = U M(xi)): the U recursion
      and 30 per half element for the four generator sums (A_u 4, A1 6,
      A2 and A3 10 each); plain dB: 3 x 4 per (Z element, role)."""
    c = so4_counts(twol)
    z = 10 * c["zterms"] + 4 * c["zpairs"]
    hu = c["nhalf"] - 1
    u = 18 * hu
    du = 102 * hu
    adj_contract = 16 * c["nhalf"]
    sweep = u + 30 * c["nhalf"]
    return dict(z_per_atom=z, u_per_nbr=u, du_per_nbr=du + u, adj_contract_per_nbr=adj_contract,
                adj_sweep_per_nbr=sweep, adjoint_per_env=z + k_mean * (u + sweep), counts=c)


# ------------------------------------------------------------------ GPU
def build_case(reps=8, sigma=0.08, seed=0):
    from pyxtal_ff_amd import workloads
    return workloads.diamond_si(reps=reps, sigma=sigma, seed=seed)


class Timers:
    """HIP-event timing of each library call (torch's current stream is the
    stream every pxff call is launched on)."""

    def __init__(self):
        self.acc = {}

    def wrap(self, mod, names):
        import torch
        for n in names:
            f = getattr(mod, n)

            def g(*a, _f=f, _n=n, **k):
                s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                s.record()
                out = _f(*a, **k)
                e.record()
                self.acc.setdefault(_n, []).append((s, e))
                return out
            setattr(mod, n, g)
            setattr(self, "_orig_" + n, f)

    def unwrap(self, mod, names):
        for n in names:
            setattr(mod, n, getattr(self, "_orig_" + n))

    def summary(self):
        import torch
        torch.cuda.synchronize()
        return {k: float(np.mean([s.elapsed_time(e) for s, e in v])) for k, v in self.acc.items()}


ADJ_KERNELS = ("ui", "z", "bi", "network", "yi", "deidrj", "gather")
PLAIN_KERNELS = ("dbdr", "rows", "dxdr_contract")
# Python wrapper -> the one kernel it launches (rocprof names)
KERNEL_SYMBOL = {"ui": "so4_accum_kernel", "z": "so4_z_tile_kernel", "bi": "so4_b_kernel",
                 "zb": "so4_z_tile_kernel+so4_brow_sum_kernel",
                 "network": "mlp_energy_grad_kernel", "yi": "so4_yi_kernel",
                 "deidrj": "so4_deidrj_kernel", "gather": "so4_gather_kernel",
                 "dbdr": "so4_dbdr_centre_kernel", "dbdr_atom_major_z": "so4_z_atom_major_kernel+so4_dbdr_centre_kernel",
                 "rows": "so4_rows_kernel", "dxdr_contract": "dxdr_contract_kernel"}


_SPIN_CYCLES_PER_MS = None


def gpu_time_ms(fn, reps=5):
    """Mean GPU duration of fn's launches.  After a warm-up call a GPU spin
    (torch.cuda._sleep) is queued ahead of the bracketing events, so every
    timed launch is enqueued while the GPU is still busy and host time between
    the calls is not counted (back-to-back Python calls of a short kernel
    otherwise leave the GPU idle between launches)."""
    import torch
    global _SPIN_CYCLES_PER_MS
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    if _SPIN_CYCLES_PER_MS is None:
        a.record()
        torch.cuda._sleep(1 << 20)
        b.record()
        b.synchronize()
        _SPIN_CYCLES_PER_MS = (1 << 20) / max(a.elapsed_time(b), 1e-3)
    # about 2 ms of spin, plus 1 ms per call of host time
    torch.cuda._sleep(int(_SPIN_CYCLES_PER_MS * (2.0 + reps)))
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / reps


def isolated_kernel_ms(dso4, potential, model, species_atom, nl, cfg, w, reps=5):
    """Mean duration of each pxff kernel of both paths, from reps
    back-to-back launches between two HIP events on the launch stream
    (gpu_time_ms)."""

    def t(fn):
        return gpu_time_ms(fn, reps)

    out = {}
    u = dso4.ui(nl, cfg, w, w)
    out["ui"] = t(lambda: dso4.ui(nl, cfg, w, w))
    z = dso4.z(nl, cfg, u)
    out["z"] = t(lambda: dso4.z(nl, cfg, u))
    b = dso4.bi(nl, cfg, u, z)
    out["bi"] = t(lambda: dso4.bi(nl, cfg, u, z))
    e, beta = model.energy_grad(b, species_atom)
    out["network"] = t(lambda: model.energy_grad(b, species_atom))
    y = dso4.yi(nl, cfg, z, beta)
    out["yi"] = t(lambda: dso4.yi(nl, cfg, z, beta))
    dedr = dso4.deidrj(nl, cfg, w, y)
    out["deidrj"] = t(lambda: dso4.deidrj(nl, cfg, w, y))
    out["gather"] = t(lambda: dso4.gather(nl, dedr))
    db = dso4.dbdr(nl, cfg, w, z)
    # the plain path's dB/dr (per-centre kernel over the tile-major Z), and
    # the alternative over an atom-major copy of Z (copy included)
    out["dbdr"] = t(lambda: dso4.dbdr(nl, cfg, w, z))
    out["dbdr_atom_major_z"] = t(lambda: dso4.dbdr(nl, cfg, w, z, atom_major=True))
    dx, _ = dso4.rows(nl, cfg.nb, db, stress=False)
    out["rows"] = t(lambda: dso4.rows(nl, cfg.nb, db, stress=False))
    out["dxdr_contract"] = t(lambda: dso4.dxdr_contract(nl, cfg.nb, beta, dx))
    return out


def copy_calibration(nbytes=1 << 28, reps=3):
    """A plain device copy of known size (for scaling FETCH_SIZE / WRITE_SIZE
    counters to bytes): reads and writes nbytes per launch."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    return nbytes


def run_gpu(steps, warmup):
    import torch
    from pyxtal_ff_amd import adjoint, potential
    from pyxtal_ff_amd import so4 as dso4
    from pyxtal_ff_amd.neighbors import NeighborListBuilder

    s = build_case()
    n = len(s)
    twol, rc = 8, 4.9
    cfg = dso4.So4Config(twol, rc, 0.99363, "cosine")
    dev = torch.device("cuda")
    pos_dev = torch.as_tensor(s.positions, dtype=torch.float64, device=dev)
    w = torch.ones(n, dtype=torch.float64, device=dev)
    model = potential.NNPotential(["Si"], cfg.nb, (30, 30), seed=0).to(dev)
    species = potential.species_index(s.symbols, ["Si"], dev)
    species_atom = potential.species_atom(s.symbols, ["Si"], dev)

    # structure metadata (cell, pbc, bin grid) is set up once, as in an MD
    # run; every step rebuilds the neighbour list from the HBM positions
    builder = NeighborListBuilder([n], [s.cell], [s.pbc], rc)

    def nlist():
        return builder.build(pos_dev)

    # launch parameters picked on this workload
    from pyxtal_ff_amd import autotune
    nl0 = nlist()
    tuned = autotune.tune_so4(nl0, cfg, w, w, use_cache=False)

    def step_adjoint():
        # adjoint kernels called directly, network forward + backward fused
        nl = nlist()
        st = adjoint.SO4State(nl, cfg, w, w)
        return potential.energy_forces_fused(nl, st, model, species_atom)

    def step_adjoint_autograd():
        # the same kernels behind torch.autograd, torch network
        nl = nlist()
        st = adjoint.SO4State(nl, cfg, w, w)
        return potential.energy_forces_adjoint(nl, st, model, species)

    def step_plain():
        nl = nlist()
        return potential.energy_forces_plain(nl, cfg, w, w, model, species)

    # per-kernel durations: each kernel launched back to back on the same
    # inputs (the queue never drains, so no host launch gaps are counted);
    # they name the dominant kernel of the adjoint step
    nl = nlist()
    ktimes = isolated_kernel_ms(dso4, potential, model, species_atom, nl, cfg, w)
    dom = max((k for k in ADJ_KERNELS if k in ktimes), key=ktimes.get)

    def timed(fn, k, wu, on_start=None):
        for _ in range(wu):
            fn()
        torch.cuda.synchronize()
        if on_start:
            on_start()
        t0 = time.perf_counter()
        for _ in range(k):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k, out

    # timed adjoint steps; HIP events on the launch stream bracket the
    # dominant kernel's launches inside the timed region
    td = Timers()
    owner = potential.NNPotential if dom == "network" else dso4
    # the function of the fused step that launches the dominant kernel (the
    # step calls zb: the Z kernel and the short B row-partial sum)
    name = {"network": "energy_grad", "z": "zb"}.get(dom, dom)
    td.wrap(owner, [name])
    t_adj, out_a = timed(step_adjoint, steps, warmup, on_start=td.acc.clear)
    dom_ms = td.summary()[name]
    td.unwrap(owner, [name])
    # the same E+F step captured once as a HIP graph and replayed; every
    # replay waits for its image count (the capacity check), as an MD
    # driver would before using the forces
    gstep = potential.FusedEFGraph(builder, cfg, (w, w), model, species_atom, pos_dev)
    t_graph, out_gr = timed(gstep, steps, warmup)
    out_gr = {k: v.clone() for k, v in out_gr.items()}
    t_pl, out_p = timed(step_plain, max(1, steps), max(1, warmup))
    t_ag, out_g = timed(step_adjoint_autograd, max(1, steps), max(1, warmup))
    ferr = float(torch.linalg.norm(out_a["forces"] - out_p["forces"]) / torch.linalg.norm(out_p["forces"]))
    ferr_ag = float(torch.linalg.norm(out_g["forces"] - out_a["forces"]) / torch.linalg.norm(out_a["forces"]))
    ferr_graph = float((out_gr["forces"] - out_a["forces"]).abs().max())

    nl = nlist()
    tnl = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nlist()
        torch.cuda.synchronize()
        tnl.append(time.perf_counter() - t0)
    kmean = nl.nimg / n
    return dict(n=n, t_adj=t_adj, t_graph=t_graph, ferr_graph=ferr_graph, t_plain=t_pl, t_autograd=t_ag, ferr=ferr,
                ferr_autograd=ferr_ag, ktimes=ktimes,
                dom=dom, dom_ms=dom_ms,
                t_nlist=float(np.median(tnl)), nimg=nl.nimg, nrows=nl.nrows, kmean=kmean, tuned=tuned)


def run_batched(steps=3):
    """1024 periodic Si/O structures (8-256 atoms, O fraction 0-0.67,
    0.066 /A^3, +-5 % strain, 1.5 A minimum distance), SNAP 2J=8, rc 4.9,
    weights Si 1.0 / O 0.8: adjoint E+F+virial of the whole batch per step."""
    import torch
    from pyxtal_ff_amd import batch as bt, potential, workloads
    structs = workloads.random_sio_batch(1024, seed=0)
    b = bt.StructureBatch(structs, 4.9)
    d = bt.BatchedSNAP(b, {"Si": 1.0, "O": 0.8}, lmax=4, rcut=4.9)
    el = ["O", "Si"]
    model = potential.NNPotential(el, d.nfeat, (30, 30), seed=0).cuda()
    sp = potential.species_index(b.symbols, el, "cuda")

    def step():
        b.nl = b.builder.build(b.pos)
        d.state.nl = b.nl
        return bt.energy_forces_virial(d, model, sp)
    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    return {"structures": b.nstruct, "atoms": b.natoms, "images": b.nl.nimg, "ms_per_step": t * 1e3,
            "atom_env_per_s": b.natoms / t}


def run_training(steps=10):
    """256 structures of the batched set; per-element 55-30-30-1 tanh MLP;
    labels from a seeded teacher on the materialised dxdr; the reference's
    calculate_loss (force coefficient 0.1, stress coefficient 0.01); one step
    = network + forces and stress + their backward + Adam, eager and as a
    captured HIP graph, with the forces through the adjoint (its tangent in
    the backward) and through the dxdr / rdxdr materialised once for the
    fixed structures (the reference's way of training)."""
    import torch
    from pyxtal_ff_amd import batch as bt, potential, training, workloads
    structs = workloads.random_sio_batch(256, seed=0)
    b = bt.StructureBatch(structs, 4.9)
    d = bt.BatchedSNAP(b, {"Si": 1.0, "O": 0.8}, lmax=4, rcut=4.9)
    el = ["O", "Si"]
    sp = potential.species_index(b.symbols, el, "cuda")
    teacher = potential.NNPotential(el, d.nfeat, (30, 30), seed=100).cuda()
    labels = training.teacher_labels(d, teacher, sp, plain=True)
    out = {"structures": b.nstruct, "atoms": b.natoms}
    for mode in ("eager", "graph", "materialized_eager", "materialized_graph"):
        student = potential.NNPotential(el, d.nfeat, (30, 30), seed=5).cuda()
        tr = training.ForceTrainer(d, student, sp, labels, graph=mode.endswith("graph"),
                                   materialized=mode.startswith("materialized"))
        if mode.startswith("materialized"):
            out["dxdr_rows"] = int(tr.dx.shape[0])
            out["dxdr_bytes"] = int(tr.dx.numel() * 8)
        l0 = float(tr.step())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            tr.step()
        torch.cuda.synchronize()
        out[mode + "_ms_per_step"] = (time.perf_counter() - t0) / steps * 1e3
        out[mode + "_loss_first_last"] = [l0, float(tr.step())]
        del tr
    return out


def so3_flops(nmax, lmax):
    """Algorithmic FP64 flops per image of the SO3 kernels, counted from the
    reference's compute_cs / compute_dcs loop bounds (SO3.py:504-694) with
    NQ = 10 (nmax + lmax + 1) quadrature nodes (a multiply-add is 2 flops,
    complex operations as for SO4): the radial integrals I and dI are
    nmax x NQ x (lmax + 1) multiply-adds each; the expansion c = e fc I Y is
    8 flops per (n, l, m); the adjoint image term forms dc_k for the 3
    directions (I grad Y + dI rhat Y + grad(e fc) I Y, 8 flops each) and
    contracts it with conj(A) (4 flops) per (n, l, m, k)."""
    nq = 10 * (nmax + lmax + 1)
    nlm = nmax * (lmax + 1) ** 2
    return dict(radial=2 * 2 * nmax * nq * (lmax + 1), ctot=8 * nlm, image_adjoint=3 * (8 + 4) * nlm)


def so3_expansion_bytes(nimg, natoms, nmax, lmax):
    """Algorithmic HBM bytes of the SO3 expansion with the radial store: the
    (lmax + 1) 2 nmax doubles of every image's radial row written once and
    read once, the separations read by both kernels, the neighbour index and
    the reverse image (int64 each), and the per-centre expansion written
    (nmax (lmax + 1) (2 lmax + 1) complex doubles per centre)."""
    nf = (lmax + 1) * 2 * nmax
    return nimg * (2 * nf * 8 + 2 * 3 * 8 + 2 * 8) + natoms * nmax * (lmax + 1) * (2 * lmax + 1) * 16


SO3_KERNELS = {"expansion": "so3_rbin_hist_kernel+so3_rbin_scan_kernel+so3_rbin_fill_kernel+"
                            "so3_radial_tab_kernel+so3_ctot_lm_kernel",
               "power": "so3_power_kernel", "network": "mlp_mfma_kernel", "adjoint_a": "so3_adj_a_kernel",
               "deidrj": "so3_deidrj_lm_kernel", "forces_energy": "so4_gather_kernel"}


def committed_traffic_sum(symbols):
    """Sum of committed_traffic over '+'-joined kernel names (None if any is
    missing)."""
    tot = 0
    for k in symbols.split("+"):
        v = committed_traffic(k)
        if v is None:
            return None
        tot += v
    return tot


def run_so3_sio2(steps=5):
    """Amorphous SiO2, 4000 Si + 8000 O at 2.2 g/cm^3 (1.6 A minimum distance),
    SO3 nmax = lmax = 6, rc 5.0, alpha 2.0, Z weights: adjoint E+F per step
    (neighbour list rebuilt from the HBM positions; per-image expansions
    recomputed in the backward, only the radial integrals kept): the fused
    kernels called directly, the same step as one captured HIP graph, and
    the autograd variant (torch network)."""
    import torch
    from pyxtal_ff_amd import batch as bt, potential, workloads
    from pyxtal_ff_amd import so3 as dso3
    from pyxtal_ff_amd import so4 as dso4
    from pyxtal_ff_amd.neighbors import NeighborListBuilder
    s = workloads.amorphous_sio2(4000, 8000, seed=0)
    n = len(s)
    el = ["O", "Si"]
    cfg = dso3.So3Config(6, 6, 5.0, 2.0)
    pos = torch.as_tensor(s.positions, dtype=torch.float64, device="cuda")
    z = torch.as_tensor(s.numbers, dtype=torch.int64, device="cuda")
    model = potential.NNPotential(el, cfg.ncoef, (30, 30), seed=0).cuda()
    spa = potential.species_atom(s.symbols, el, "cuda")
    sp = potential.species_index(s.symbols, el, "cuda")
    builder = NeighborListBuilder([n], [s.cell], [s.pbc], 5.0)

    def fused():
        nl = builder.build(pos)
        return potential.energy_forces_fused(nl, dso3.SO3State(nl, cfg, z), model, spa)

    b = bt.StructureBatch([s], 5.0)
    d = bt.BatchedSO3(b, nmax=6, lmax=6, rcut=5.0, alpha=2.0)

    def autograd():
        b.nl = b.builder.build(b.pos)
        d.state.nl = b.nl
        return bt.energy_forces_virial(d, model, sp, virial=False)

    def timed(fn, k):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k, out

    t_fused, o_f = timed(fused, steps)
    g = potential.FusedEFGraph(builder, cfg, (z,), model, spa, pos)
    t_graph, o_g = timed(g, steps)
    t_ag, o_a = timed(autograd, max(1, steps // 2))
    ferr = float(torch.linalg.norm(o_a["forces"] - o_f["forces"]) / torch.linalg.norm(o_f["forces"]))
    geq = bool(torch.equal(o_g["forces"], o_f["forces"]))
    # per-kernel durations, each launched back to back on the same inputs
    nl = builder.build(pos)
    nimg = nl.nimg

    def t(fn, reps=5):
        return gpu_time_ms(fn, reps)

    kt = {}
    x, ct, rad = dso3.forward(nl, cfg, z, radial=True)
    # the expansion: radial integrals once per image pair, then the per-centre sums
    kt["expansion"] = t(lambda: dso3.ctot(nl, cfg, z, radial=True))
    kt["power"] = t(lambda: dso3.power(nl, cfg, ct))
    e, beta = model.energy_grad(x, spa)
    kt["network"] = t(lambda: model.energy_grad(x, spa))
    A = dso3.adjoint_a(nl, cfg, ct, beta)
    kt["adjoint_a"] = t(lambda: dso3.adjoint_a(nl, cfg, ct, beta))
    kt["deidrj"] = t(lambda: dso3.deidrj(nl, cfg, z, A, radial=rad))
    dedr = dso3.deidrj(nl, cfg, z, A, radial=rad)
    kt["forces_energy"] = t(lambda: dso4.forces_energy(nl, dedr, e))
    fl = so3_flops(6, 6)
    # the expansion (tabulated radial integrals, then the per-centre sums) is
    # priced by its HBM bytes, the adjoint image sweep by its flops
    ref_flops = (fl["radial"] + fl["ctot"]) * nimg
    roof = {"expansion": ("hbm", so3_expansion_bytes(nimg, n, 6, 6), PEAK_HBM, "GB/s", 1e9),
            "deidrj": ("valu", fl["image_adjoint"] * nimg, PEAK_FP64, "TFLOP/s", 1e12)}
    dom = max(roof, key=lambda k: kt[k])
    bound, work, peak, unit, scale = roof[dom]
    ach = work / (kt[dom] * 1e-3)
    return {"atoms": n, "images": nimg, "ms_per_step": t_fused * 1e3, "atom_env_per_s": n / t_fused,
            "graph_ms_per_step": t_graph * 1e3, "graph_atom_env_per_s": n / t_graph,
            "graph_forces_equal_eager": geq, "autograd_ms_per_step": t_ag * 1e3,
            "force_rel_diff_autograd_vs_fused": ferr, "kernel_ms_isolated": kt,
            "roofline": {"bound": bound, "kernel": SO3_KERNELS[dom], "achieved": ach / scale, "peak": peak / scale,
                         "unit": unit, "frac": ach / peak, "work_per_launch": work,
                         "traffic": committed_traffic_sum(SO3_KERNELS[dom]),
                         "kernel_ms_isolated": kt[dom],
                         "reference_work_tflops": ref_flops / (kt["expansion"] * 1e-3) / 1e12,
                         "note": "expansion: algorithmic bytes (bench.so3_expansion_bytes); deidrj: algorithmic "
                                 "flops from the reference loop bounds (bench.so3_flops). reference_work_tflops is "
                                 "the reference's quadrature + expansion flops over the expansion's time: the "
                                 "tabulated radial integrals do not execute those flops, so it is a rate of "
                                 "reference work done, not of arithmetic"}}


# ------------------------------------------------------------ CPU baseline
def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(reps_one=3, reps_all=5):
    """The oracle (C restatement of the reference's numba kernels: two passes,
    U recomputed per pair, materialised dB/dr and per-row rdxdr, stress=True
    as the reference's Database.compute / compute_descriptor call it) on the
    same 4096-atom structure, with the reference's force contraction over
    the rows: on one host core (the reference runs one structure serially)
    and on every host core (OpenMP over centres; the reference fans out over
    structures with a process pool instead)."""
    import oracle
    oracle.build()
    s = build_case()
    ncores = len(os.sched_getaffinity(0))

    def once(threads):
        t0 = time.perf_counter()
        d = oracle.snap_calculate(s.positions, s.cell, s.pbc, s.symbols, {"Si": 1.0}, lmax=4, rcut=4.9,
                                  derivative=True, stress=True, threads=threads)
        beta = np.ones_like(d["x"])
        f = np.zeros((len(s), 3))
        np.add.at(f, d["seq"][:, 1], -np.einsum("rb,rbk->rk", beta[d["seq"][:, 0]], d["dxdr"]))
        return time.perf_counter() - t0

    once(ncores)  # warm-up (tables, page faults)
    t_all = float(np.median([once(ncores) for _ in range(reps_all)]))
    t_one = float(np.median([once(1) for _ in range(reps_one)]))
    n = len(s)
    return dict(value=n / t_all, unit="atom-env/s", cores=ncores, kind="port",
                sample="full 4096-atom diamond Si 2J=8 structure: neighbour layout, descriptor, dxdr, per-row "
                       "rdxdr (stress=True) and the force contraction; median of %d runs of %.2f s on %d threads "
                       "(OpenMP over centres), %s" % (reps_all, t_all, ncores, _cpu_model()),
                one_core={"value": n / t_one, "cores": 1,
                          "sample": "same, median of %d runs of %.2f s on 1 thread" % (reps_one, t_one)})


def committed_traffic(symbol):
    """HBM bytes per launch of a kernel on this workload, from the committed
    rocprofv3 --pmc summary (profiles/*traffic*.md lines
    'TRAFFIC <kernel> fetch=<bytes> write=<bytes>'), or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.md"))):
        with open(path) as f:
            for line in f:
                m = re.match(r"\s*TRAFFIC\s+(\S+)\s+fetch=(\d+)\s+write=(\d+)", line)
                if m and m.group(1) == symbol:
                    return int(m.group(2)) + int(m.group(3))
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched / training / SO3 workloads")
    ap.add_argument("--extra", default="batched,training,so3",
                    help="comma-separated secondary workloads to run (batched, training, so3)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 256 MiB device copy (scale reference for rocprofv3 --pmc byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the workload is one GPU's share, and "
              "the multi-GPU split (spatial slabs with a ghost-atom halo for one "
              "large structure, whole structures for a batch) is not built", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        copy_calibration()
    r = run_gpu(args.steps, args.warmup)
    n = r["n"]
    adj_rate = n / r["t_adj"]
    graph_rate = n / r["t_graph"]
    plain_rate = n / r["t_plain"]
    fl = flops(8, r["kmean"])
    kt = r["ktimes"]
    dom = r["dom"]
    alg = {"z": fl["z_per_atom"] * n,
           "deidrj": fl["adj_sweep_per_nbr"] * r["nimg"],
           "ui": fl["u_per_nbr"] * r["nimg"],
           "yi": 8 * so4_counts(8)["nhalf"] * n,
           "network": 4 * (55 * 30 + 30 * 30 + 30) * n}.get(dom)
    achieved = alg / (r["dom_ms"] * 1e-3) if alg else None
    cpu = None if args.no_cpu else run_cpu_baseline()
    extra = {}
    if not args.no_extra:
        chosen = set(args.extra.split(","))
        for key, name, fn in (("batched", "batched_sio_1024", run_batched), ("training", "training_256", run_training),
                              ("so3", "so3_sio2_12000", run_so3_sio2)):
            if key not in chosen:
                continue
            try:
                extra[name] = fn()
            except Exception as e:  # a secondary workload must not hide the headline line
                traceback.print_exc(file=sys.stderr)
                extra[name] = {"error": repr(e)[:300]}
    line = {
        "metric": METRIC, "value": adj_rate, "unit": "atom-env/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["t_adj"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "diamond Si 8x8x8 (4096 atoms), sigma 0.08 A, SNAP 2J=8, rc 4.9, E+F",
                   "natoms": n, "images": r["nimg"], "seq_rows": r["nrows"],
                   "path": "neighbour list rebuilt from the HBM positions, adjoint kernels called directly, "
                           "network forward+backward fused",
                   "graph_ms_per_step": r["t_graph"] * 1e3, "graph_atom_env_per_s": graph_rate,
                   "force_max_abs_diff_graph_vs_eager": r["ferr_graph"],
                   "plain_dxdr_atom_env_per_s": plain_rate, "plain_ms_per_step": r["t_plain"] * 1e3,
                   "adjoint_over_plain": adj_rate / plain_rate, "force_rel_diff_adjoint_vs_plain": r["ferr"],
                   "adjoint_autograd_ms_per_step": r["t_autograd"] * 1e3,
                   "force_rel_diff_autograd_vs_fused": r["ferr_autograd"],
                   "nlist_ms": r["t_nlist"] * 1e3, "kernel_ms_isolated": kt, "tuned": r["tuned"], **extra},
        "roofline": {"bound": "valu", "kernel": KERNEL_SYMBOL[dom],
                     "achieved": achieved / 1e12 if achieved else None, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64 if achieved else None,
                     "traffic": committed_traffic(KERNEL_SYMBOL[dom]),
                     "kernel_ms_timed_region": r["dom_ms"],
                     "flops_per_launch": alg,
                     "note": "FP64 vector ALU roof (no MFMA; MI355X FP64 vector peak = FP64 matrix peak)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
