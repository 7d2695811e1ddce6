"""Batched steps: members of a batch advanced by one captured graph of
persistent kernels must match the same simulations run one by one."""
import sys

import numpy as np
import pytest

from vdropimpact_amd import workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


FIELDS = None


def fields(vdi):
    return (vdi.U, vdi.V, vdi.P, vdi.PG, vdi.H, vdi.HT, vdi.GP, vdi.GTAU)


@pytest.mark.parametrize("form", [0, 1])
def test_gpu_batch_matches_single(vdi, form):
    # the mixed form is active on grids with a tiled level (above 3072 unknowns)
    grids = [(64, 32), (128, 64), (256, 128)] if form == 0 else [(128, 64), (256, 128)]
    cfgs, prms = workloads.batch_members(5, grids=grids)
    say("batch vs single, form", form, [(p.m, p.n) for p in prms])
    b = vdi.Batch(prms, form)
    b.init_fields()
    singles = []
    for p in prms:
        s = vdi.Fluid2D(p)
        s.set_form(form)
        s.init_fields()
        singles.append(s)
    for step in range(12):
        b.step()
        for s, p in zip(singles, prms):
            s.step_forward(p.dt_reg)
    nb, redo = b.counters()
    say("  batched steps", nb, "redone members", redo)
    assert nb >= 8
    for k, (s, p) in enumerate(zip(singles, prms)):
        m = b.sim(k)
        for aid in fields(vdi):
            e = rel(m.get(aid), s.get(aid))
            assert e < 1e-9, (k, aid, e)
        assert m.time == s.time
        assert m.stats().steps == 12
    g, busy = b.step_times()
    say("  graph ms", g, "busy ms", busy)
    assert g > 0 and np.all(busy > 0)
    for s in singles:
        s.close()
    b.close()


def test_gpu_batch_autotune_keeps_results(vdi):
    grids = [(64, 32), (128, 64)]
    cfgs, prms = workloads.batch_members(3, grids=grids)
    say("batch autotune", [(p.m, p.n) for p in prms])
    a = vdi.Batch(prms, 0)
    c = vdi.Batch(prms, 0)
    a.init_fields()
    c.init_fields()
    for _ in range(3):
        a.step()
        c.step()
    a.autotune(2)
    say("  tiles", a.tiles()[:12])
    for _ in range(6):
        a.step()
        c.step()
    for k in range(len(prms)):
        for aid in fields(vdi):
            assert rel(a.sim(k).get(aid), c.sim(k).get(aid)) < 1e-9
    a.close()
    c.close()
