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


def oracle_end_step(oracle, p, limit):
    """The step at which the oracle's run of p stops on the gas layer's
    "Too many Newton steps" (gas_layer_solve.cc:204), or None."""
    o = oracle.OracleSim(p, rel_tol=1e-10)
    o.init_fields()
    try:
        for k in range(limit):
            try:
                o.step_forward(p.dt_reg)
            except RuntimeError as e:
                assert "status 2" in str(e)   # the Newton loop's own limit
                return k
        return None
    finally:
        o.close()


def test_gpu_batch_member_end_and_restart(vdi, oracle):
    """Members 1 and 48 of the 64-simulation mix (256x128, 0.76 and 0.71 m/s):
    member 1's gas film closes and its run ends with the gas layer's Newton
    error at the oracle's step (within one step: the count at the threshold
    step depends on the last digits); member 48 runs on.  The ended member,
    restarted, matches a newly created simulation step for step, and member
    48 still matches its single run."""
    cfgs, prms = workloads.batch_members(64)
    pa, pb = prms[1], prms[48]
    assert (pa.m, pa.n) == (256, 128) and (pb.m, pb.n) == (256, 128)
    end = oracle_end_step(oracle, pa, 420)
    say("oracle run of member 1 ends at step", end)
    assert end is not None and end > 300
    b = vdi.Batch([pa, pb], vdi.FORM_FP64)
    b.init_fields()
    sb = vdi.Fluid2D(pb)
    sb.init_fields()
    ended = None
    for k in range(end + 3):
        try:
            b.step()
        except vdi.VdiError as e:
            assert e.status == vdi.ERR_NEWTON_NOCONV
            st = b.member_status()
            assert st[0] == vdi.ERR_NEWTON_NOCONV and st[1] == vdi.OK
            ended = k
        sb.step_forward(pb.dt_reg)
        if ended is not None:
            break
    say("batched member 1 ends at step", ended, "batched steps", b.counters())
    assert ended is not None and abs(ended - end) <= 1
    # an ended member must be restarted before the batch steps again
    with pytest.raises(vdi.VdiError):
        b.step()
    b.restart(0)
    assert b.sim(0).time == 0.0
    fresh = vdi.Fluid2D(pa)
    fresh.init_fields()
    for k in range(6):
        b.step()
        fresh.step_forward(pa.dt_reg)
        sb.step_forward(pb.dt_reg)
        for aid in fields(vdi):
            e0 = rel(b.sim(0).get(aid), fresh.get(aid))
            e1 = rel(b.sim(1).get(aid), sb.get(aid))
            assert e0 < 1e-9 and e1 < 1e-9, (k, aid, e0, e1)
        assert b.sim(0).time == fresh.time and b.sim(1).time == sb.time
    fresh.close()
    sb.close()
    b.close()
