"""The benchmark configurations at their own sizes on the GPU, against the
oracle:

- the advection passes alone (the Godunov predictor with the MAC right-hand
  side, fluid_2d.cc:105-232, then the MAC correction and advective update,
  fluid_2d.cc:236-261) on the 64-mode Fourier field at 1024 x 1024 and
  2048 x 1024 (the solves of that operator-level configuration are in
  test_gpu_operator.py);
- the batch of 64 drops (grids drawn from 256 x 128, 512 x 256 and
  1024 x 512, viscosity, speed and radius drawn per drop) in the mixed form:
  every member against the same drop run alone, and one member of each grid
  size against the FP64 oracle;
- the learned initial guess at 1024 x 512: trained on pairs the solver
  records at that size, the fields then agree with the plain FP64 path.

Parity bar: 1e-9 relative L2 per field (1e-12 for single passes, which
involve no solve).
"""
import sys

import numpy as np
import pytest

import test_gpu_parity as tp
from vdropimpact_amd import workloads
from vdropimpact_amd.params import PhysicalConfig, VdiParams

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def advection_params(m, n):
    return VdiParams(m=m, n=n, x_sym=1, implicit_visc=1, ax=0.0, bx=1.0, ay=0.0, by=n / m, rho=1.0, mu=1e-3,
                     dt_reg=0.2 / m, alpha=1 / 1.4, mug=1.8e-5, Pamb=1e5, h0=5e-6, R=1e-3, V=0.5, sigma=0.02)


@pytest.mark.parametrize("m,n", [(1024, 1024), (2048, 1024)])
def test_gpu_advection_passes_match_oracle(vdi, oracle, m, n):
    say("advection passes", m, n)
    p = advection_params(m, n)
    u, v, pn, z = workloads.advection_state(m, n)
    gpu = vdi.Fluid2D(p)
    orc = oracle.OracleSim(p)
    for s, ids in ((gpu, (vdi.U, vdi.V, vdi.P)), (orc, (oracle.U, oracle.V, oracle.P))):
        for aid, arr in zip(ids, (u, v, pn)):
            s.set(aid, arr)
    dt = p.dt_reg
    gpu.run_pass(vdi.PASS_GODUNOV, dt)
    for k in (oracle.PASS_K1, oracle.PASS_K2, oracle.PASS_K3, oracle.PASS_K4, oracle.PASS_K5, oracle.PASS_K6):
        orc.run_pass(k, dt)
    # faces the later passes read: vertical faces 0..m of rows 0..n-1,
    # horizontal faces 0..n of columns 0..m-1
    R, C = slice(2, n + 2), slice(2, m + 2)
    worst = 0.0
    for name, ga, oa, sl in (("ul", vdi.UL, oracle.UL, (R, slice(2, m + 3))), ("vl", vdi.VL, oracle.VL, (R, slice(2, m + 3))),
                             ("vd", vdi.VD, oracle.VD, (slice(2, n + 3), C)), ("ud", vdi.UD, oracle.UD, (slice(2, n + 3), C)),
                             ("c0", vdi.C0, oracle.C0, (R, C)), ("c1", vdi.C1, oracle.C1, (R, C))):
        e = rel(gpu.padded(ga)[sl], orc.padded(oa)[sl])
        worst = max(worst, e)
        assert e < 1e-12, (name, e)
    src_g, src_o = gpu.get(vdi.SRC), orc.get(oracle.SRC)
    e = rel(src_g[:m * n + m], src_o[:m * n + m])
    assert e < 1e-12, ("src", e)
    gpu.set(vdi.Z_MAC, z.ravel())
    orc.set(oracle.Z_MAC, z.ravel())
    gpu.run_pass(vdi.PASS_CORRECT_ADVECT, dt)
    orc.run_pass(oracle.PASS_K7)
    orc.run_pass(oracle.PASS_K8, dt)
    for name, ga, oa in (("c0", vdi.C0, oracle.C0), ("c1", vdi.C1, oracle.C1)):
        e = rel(gpu.padded(ga)[R, C], orc.padded(oa)[R, C])
        worst = max(worst, e)
        assert e < 1e-12, (name, e)
    say("  worst rel L2", worst)
    gpu.close()
    orc.close()


def test_gpu_batch_of_64_matches_single_runs_and_oracle(vdi, oracle):
    """The batched workload as configured (64 drops, grids uniform over the
    three sizes, parameters drawn per drop, mixed form); every solve to
    1e-13 so that members compare with the FP64 oracle at the parity bar."""
    cfgs, prms = workloads.batch_members(64)
    sizes = [(p.m, p.n) for p in prms]
    say("batch of 64:", {s: sizes.count(s) for s in sorted(set(sizes))})
    b = vdi.Batch(prms, vdi.FORM_MIXED)
    for k in range(len(prms)):
        b.sim(k).set_solver(rel_tol=1e-13)
    b.init_fields()
    singles = []
    for p in prms:
        s = vdi.Fluid2D(p, rel_tol=1e-13)
        s.set_form(vdi.FORM_MIXED)
        s.init_fields()
        singles.append(s)
    pick = {}
    for k, sz in enumerate(sizes):
        pick.setdefault(sz, k)
    orcs = {}
    for sz, k in pick.items():
        o = oracle.OracleSim(prms[k], rel_tol=1e-13)
        for q in range(4):
            o.set_schedule(q, *singles[k].schedule(q))
        o.init_fields()
        orcs[k] = o
    steps = 8
    for step in range(steps):
        b.step()
        for s, p in zip(singles, prms):
            s.step_forward(p.dt_reg)
        for k, o in orcs.items():
            o.step_forward(prms[k].dt_reg)
    nb, redo = b.counters()
    say("  batched steps", nb, "redone members", redo)
    assert nb >= steps - 3
    worst = 0.0
    for k, (s, p) in enumerate(zip(singles, prms)):
        mem = b.sim(k)
        assert mem.time == s.time and mem.stats().steps == steps
        for aid in (vdi.U, vdi.V, vdi.P, vdi.PG, vdi.H, vdi.GP):
            e = rel(mem.get(aid), s.get(aid))
            worst = max(worst, e)
            assert e < 1e-9, (k, sizes[k], aid, e)
    say("  worst member vs single", worst)
    for k, o in orcs.items():
        w = tp.compare(b.sim(k), o, f"member {k} {sizes[k]}")
        say("  member", k, sizes[k], "vs oracle, worst", w)
    spans = b.member_spans()
    g, _busy = b.step_times()
    say("  batched step ms", g, "member span ms p50/p99", np.percentile(spans, 50), np.percentile(spans, 99))
    assert np.all(spans > 0) and np.all(spans <= g * 1.05)
    for s in singles:
        s.close()
    for o in orcs.values():
        o.close()
    b.close()


def test_gpu_learned_guess_at_1024x512(vdi):
    """The learned guess at the metric grid: pairs recorded at 1024 x 512,
    training on the GPU, then the network's guess (with the residual
    fallback) in the MAC solve; the fields equal the plain FP64 path's to the
    solve tolerance.  The cycles it saves or costs are reported."""
    p = PhysicalConfig(m=1024, n=512).params()
    a, b = vdi.Fluid2D(p), vdi.Fluid2D(p)
    a.init_fields()
    b.init_fields()
    a.guess_attach(capacity=32, seed=5)
    a.guess_mode(1)
    for _ in range(34):
        a.step_forward(p.dt_reg)
        b.step_forward(p.dt_reg)
    assert a.guess_stats()["pairs"] == 32
    loss = a.guess_train(epochs=8, batch=4, lr=1e-3)
    say("training loss", loss)
    assert np.isfinite(loss)
    a.guess_mode(2)
    a.set_graphs(False)
    b.set_graphs(False)
    a.reset_stats()
    b.reset_stats()
    K = 12
    for _ in range(K):
        a.step_forward(p.dt_reg)
        b.step_forward(p.dt_reg)
    st = a.guess_stats()
    ca, cb = a.stats().total_cycles[0] / K, b.stats().total_cycles[0] / K
    say(f"MAC V-cycles per step: {ca:.3f} with the network, {cb:.3f} with extrapolation; "
        f"network guess kept {st['accepted']} of {K}")
    assert st["accepted"] + st["rejected"] == K
    for aid in (vdi.U, vdi.V, vdi.P, vdi.PG, vdi.H):
        e = rel(a.get(aid), b.get(aid))
        assert e < 1e-8, (aid, e)
    a.close()
    b.close()
