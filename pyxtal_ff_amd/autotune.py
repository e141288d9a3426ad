"""Autotuner for the launch parameters of the hot kernels.

Knobs (``pxff_tune_set``, include/pxff.h):
  * Z-row chunks per 64-atom tile of the Z kernel (wavefronts per tile),
  * persistent-grid size of the per-centre kernels (wavefronts per CU),
  * how the per-centre kernels sum over images (register sums reduced once
    per centre, per-group LDS partials, LDS atomic adds, or LDS partials
    shared by two or three groups),
  * wavefronts per workgroup of the per-image adjoint kernel.
Each knob is timed on the caller's own workload with HIP events recorded on
the stream the kernels run on (torch's current stream), median of a few
launches; the best value is kept in the process and cached on disk, keyed by
the kernel, 2J and a power-of-two bucket of the problem size.
"""
import json
import os

import torch

from . import _native as nat
from . import so4 as dso4

Z_CHUNKS, ACCUM_WAVES_PER_CU, IMAGE_WAVES_PER_BLOCK, ACCUM_LDSADD = 0, 1, 2, 3
CANDIDATES = {Z_CHUNKS: (4, 8, 16, 24, 32, 48, 64), ACCUM_WAVES_PER_CU: (4, 8, 16, 32, 64),
              IMAGE_WAVES_PER_BLOCK: (1, 2, 4), ACCUM_LDSADD: (4, 0, 1, 2, 3)}
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "pyxtal_ff_amd", "tune.json")


def set_knob(knob, value):
    nat.check(nat.lib().pxff_tune_set(int(knob), int(value)), "pxff_tune_set")


def get_knob(knob):
    return int(nat.lib().pxff_tune_get(int(knob)))


def _time(fn, reps=5):
    ts = []
    for _ in range(reps + 1):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b))
    ts = sorted(ts[1:])
    return ts[len(ts) // 2]


def _bucket(n):
    return 1 << max(0, int(n - 1).bit_length())


def _load_cache():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(c):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def tune_so4(nl, cfg, atom_w, site_w, work_order=None, use_cache=True, verbose=False):
    """Pick the three knobs for this SO4 workload; returns {knob: value}."""
    key = "so4:2J=%d:atoms=%d:images=%d" % (cfg.twol, _bucket(nl.natoms), _bucket(nl.nimg))
    cache = _load_cache() if use_cache else {}
    if key in cache:
        best = {int(k): v for k, v in cache[key].items()}
        for k, v in best.items():
            set_knob(k, v)
        return best
    best = {}
    # U_tot (per-centre persistent grid)
    res = {}
    for mode in CANDIDATES[ACCUM_LDSADD]:
        set_knob(ACCUM_LDSADD, mode)
        for v in CANDIDATES[ACCUM_WAVES_PER_CU]:
            set_knob(ACCUM_WAVES_PER_CU, v)
            res[(mode, v)] = _time(lambda: dso4.ui(nl, cfg, atom_w, site_w, work_order))
    best[ACCUM_LDSADD], best[ACCUM_WAVES_PER_CU] = min(res, key=res.get)
    set_knob(ACCUM_LDSADD, best[ACCUM_LDSADD])
    set_knob(ACCUM_WAVES_PER_CU, best[ACCUM_WAVES_PER_CU])
    if verbose:
        print("accum (ldsadd, waves/CU)", res)
    u = dso4.ui(nl, cfg, atom_w, site_w, work_order)
    res = {}
    for v in CANDIDATES[Z_CHUNKS]:
        set_knob(Z_CHUNKS, v)
        res[v] = _time(lambda: dso4.z(nl, cfg, u))
    best[Z_CHUNKS] = min(res, key=res.get)
    set_knob(Z_CHUNKS, best[Z_CHUNKS])
    if verbose:
        print("Z chunks", res)
    z = dso4.z(nl, cfg, u)
    beta = torch.ones((nl.natoms, cfg.nb), dtype=torch.float64, device=nl.pos.device)
    y = dso4.yi(nl, cfg, z, beta)
    res = {}
    for v in CANDIDATES[IMAGE_WAVES_PER_BLOCK]:
        set_knob(IMAGE_WAVES_PER_BLOCK, v)
        res[v] = _time(lambda: dso4.deidrj(nl, cfg, atom_w, y))
    best[IMAGE_WAVES_PER_BLOCK] = min(res, key=res.get)
    set_knob(IMAGE_WAVES_PER_BLOCK, best[IMAGE_WAVES_PER_BLOCK])
    if verbose:
        print("deidrj waves/block", res)
    if use_cache:
        cache[key] = {str(k): v for k, v in best.items()}
        _save_cache(cache)
    return best
