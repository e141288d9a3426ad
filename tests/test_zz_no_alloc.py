"""The no-allocation tangent (pxff_so4_tangent_noalloc, workspace in the
library's static device array) against the allocating pxff_so4_tangent
through the C ABI: two calls in a row, the second on a smaller structure,
bit for bit; and an oversized input refused by the host-side size check
before any launch."""
import numpy as np
import pytest
import torch

from pyxtal_ff_amd import _native as nat
from pyxtal_ff_amd import so4 as dso4
from pyxtal_ff_amd import workloads
from pyxtal_ff_amd.neighbors import DeviceNeighborList

pytestmark = pytest.mark.gpu


def _case(reps, seed):
    s = workloads.diamond_si(reps=reps, sigma=0.06, seed=seed)
    nl = DeviceNeighborList([s.positions], [s.cell], [s.pbc], 4.9).resolve()
    cfg = dso4.So4Config(8, 4.9, 0.99363)
    w = torch.ones(len(s), dtype=torch.float64, device="cuda")
    _, z, _ = dso4.forward(nl, cfg, w, w)
    g = torch.Generator(device="cpu").manual_seed(seed)
    dr = torch.randn((nl.nimg, 3), generator=g, dtype=torch.float64).cuda()
    return nl, cfg, w, z, dr


def test_noalloc_tangent_equals_allocating_tangent():
    for reps, seed in ((4, 1), (2, 2)):  # 512 atoms, then 64
        nl, cfg, w, z, dr = _case(reps, seed)
        print("noalloc tangent", nl.natoms, nl.nimg, flush=True)
        ref = dso4.tangent(nl, cfg, w, z, dr)  # also builds the 2J tables on this device
        db = torch.full((nl.natoms, cfg.nb), float("nan"), dtype=torch.float64, device="cuda")
        out = dso4.tangent_noalloc(nl, cfg, w, z, dr, db)
        torch.cuda.synchronize()
        assert torch.isfinite(out).all()
        assert torch.equal(out, ref)


def test_noalloc_tangent_refuses_inputs_beyond_its_workspace():
    nl, cfg, w, z, dr = _case(2, 3)
    dso4.tangent(nl, cfg, w, z, dr)
    natoms = (32 << 20) // (cfg.nhalf * 16) + 64  # one tile more than the 32 MiB workspace holds
    p = nat.ptr
    db = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = nat.lib().pxff_so4_tangent_noalloc(cfg.params, natoms, p(nl.img_start), p(nl.img_j), p(nl.img_r), p(w),
                                            p(dr), p(z), p(db), nat.stream())
    assert rc == -1 and "workspace" in nat.last_error()
