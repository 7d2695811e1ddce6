"""The drop-in boundary beyond the whole step, on the GPU through the C ABI:
the solver entry points mgs_mac/mgs_fem::solve_v_cycle(t) and
visco_impl::setup_and_solve(vfx, vfy, t) called in sequence against the
oracle's own sequence (time extrapolation after two solves, the duplicate-time
warning, re-setup on a changed coupling), the fixed-budget solve a captured
step enqueues, captured steps over changing time steps, and rejected
schedules.

Parity bar: <= 1e-9 relative L2 once both sides have converged to the same
residual (solves here run to 1e-12 relative residual on both sides).
"""
import sys

import numpy as np
import pytest

import test_gpu_parity as tp
from vdropimpact_amd.params import PhysicalConfig

pytestmark = pytest.mark.gpu

PARITY = 1e-9


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def smooth_field(rng, m, n, k):
    """A smooth right-hand side that drifts with k (so that the time
    extrapolation of the previous solutions is a useful guess)."""
    x = (np.arange(m) + 0.5) / m
    y = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, y)
    a = rng.standard_normal(6)
    return (np.cos(np.pi * (1 + a[0] * 0.01) * X) * np.sin(np.pi * 0.5 * Y) * (1 + 0.05 * k)
            + 0.3 * np.cos(3 * np.pi * X + 0.02 * k) * np.cos(2.5 * np.pi * Y) + 0.01 * a[1])


def pair(vdi, oracle, m=128, n=64, tol=1e-12):
    cfg = PhysicalConfig(m=m, n=n)
    p = cfg.params()
    gpu = vdi.Fluid2D(p, rel_tol=tol)
    orc = oracle.OracleSim(p, rel_tol=tol)
    for k in range(4):
        orc.set_schedule(k, *gpu.schedule(k))
    return p, gpu, orc


@pytest.mark.parametrize("which", ["mac", "fem"])
def test_gpu_solve_v_cycle_sequence_matches_oracle(vdi, oracle, which, capfd):
    """Five solves at distinct times (the third on extrapolates from the last
    two solutions, mgs_common.cc:37-44), then two at a repeated time, after
    which the reference warns and keeps the last solution as the guess
    (mgs_common.cc:32-35)."""
    say("solve_v_cycle sequence", which)
    p, gpu, orc = pair(vdi, oracle)
    m, n = p.m, p.n
    kind_g = vdi.MAC if which == "mac" else vdi.FEM
    kind_o = oracle.KIND_MAC if which == "mac" else oracle.KIND_FEM
    zid = vdi.Z_MAC if which == "mac" else vdi.Z_FEM
    M, N = (m, n) if which == "mac" else (m + 1, n + 1)
    rng = np.random.default_rng(11)
    times = [1e-6, 2e-6, 3.5e-6, 5e-6, 6e-6]
    for k, t in enumerate(times):
        b = smooth_field(rng, M, N, k).ravel()
        src = np.zeros((m + 1) * (n + 1))
        src[:M * N] = b
        gpu.set(vdi.SRC, src)
        orc.set(oracle.SRC, src)
        cg = gpu.mg_solve_v_cycle(kind_g, t)
        co = orc.solve_v_cycle(kind_o, t)
        e = rel(gpu.get(zid), orc.get(zid))
        say(f"  t={t:.2e} cycles gpu {cg} oracle {co} rel {e:.2e}")
        assert e < PARITY
        assert abs(cg - co) <= 1
    # a repeated time: the history gets two equal times and every later
    # call warns and solves from the last solution
    capfd.readouterr()
    for t in (7e-6, 7e-6, 8e-6, 9e-6):
        b = smooth_field(rng, M, N, 9).ravel()
        src = np.zeros((m + 1) * (n + 1))
        src[:M * N] = b
        gpu.set(vdi.SRC, src)
        orc.set(oracle.SRC, src)
        gpu.mg_solve_v_cycle(kind_g, t)
        orc.solve_v_cycle(kind_o, t)
        assert rel(gpu.get(zid), orc.get(zid)) < PARITY
    err = capfd.readouterr().err
    assert err.count("duplicate times") >= 4   # two calls on each side
    gpu.close()
    orc.close()


@pytest.mark.parametrize("which", ["vx", "vy"])
def test_gpu_visco_setup_and_solve_sequence_matches_oracle(vdi, oracle, which):
    """Calls with one coupling, a changed coupling (the hierarchy is set up
    again, visco_impl.cc:26-30) and the first coupling again, each against
    the oracle's visco_impl::setup_and_solve sequence on the shared src."""
    say("visco setup_and_solve sequence", which)
    p, gpu, orc = pair(vdi, oracle)
    m, n = p.m, p.n
    kind_g = vdi.VX if which == "vx" else vdi.VY
    kind_o = oracle.KIND_VX if which == "vx" else oracle.KIND_VY
    zid = vdi.Z_VX if which == "vx" else vdi.Z_VY
    rng = np.random.default_rng(23)
    calls = [(0.13, 0.11, 1e-6), (0.13, 0.11, 2e-6), (0.2, 0.17, 3e-6), (0.2, 0.17, 4e-6), (0.13, 0.11, 5e-6)]
    for k, (vfx, vfy, t) in enumerate(calls):
        src = np.zeros((m + 1) * (n + 1))
        src[:m * n] = smooth_field(rng, m, n, k).ravel()
        gpu.set(vdi.SRC, src)
        orc.set(oracle.SRC, src)
        cg = gpu.visco_setup_and_solve(kind_g, vfx, vfy, t)
        co = orc.visco_setup_and_solve(kind_o, vfx, vfy, t)
        e = rel(gpu.get(zid), orc.get(zid))
        say(f"  vfx={vfx} t={t:.1e} cycles gpu {cg} oracle {co} rel {e:.2e}")
        assert e < PARITY
        assert abs(cg - co) <= 1
    gpu.close()
    orc.close()


@pytest.mark.parametrize("kind,m,n", [(0, 2048, 1024), (1, 2049, 1025)])
def test_gpu_fixed_budget_zero_rhs_gives_zero(vdi, kind, m, n):
    """A captured step's solve (fixed budget, the start's norms finished by
    the first level-0 down kernel) with a zero right-hand side: every
    workgroup, also one dispatched after workgroup 0 raised the flag, zeroes
    its tile, so z is exactly zero."""
    say("fixed budget, zero rhs", kind, m, n)
    rng = np.random.default_rng(5)
    g = vdi.MGSystem(kind, m, n, 1.0, 1.0, 1)
    g.set_b(np.zeros(m * n))
    for budget in (1, 3):
        g.set_z(rng.standard_normal(m * n))
        done, cyc = g.solve_fixed(1e-10, budget)
        assert done == 3 and cyc == 0
        assert not np.any(g.get_z())
    # a nonzero right-hand side: the fixed-budget solve gives the plain
    # V-cycles' iterate bit for bit and reports convergence
    b = rng.standard_normal(m * n)
    g.set_b(b)
    g.set_z(np.zeros(m * n))
    c = g.solve(1e-10, 100)
    z1 = g.get_z()
    g.set_z(np.zeros(m * n))
    done, cyc = g.solve_fixed(1e-10, c + 2)
    assert done == 1 and cyc == c
    np.testing.assert_array_equal(g.get_z(), z1)
    g.close()


def test_gpu_graphs_over_changing_time_steps(vdi, oracle):
    """Captured steps over ten distinct time steps, revisited: each graph
    keeps the viscous couplings and coarse inverses it was recorded with, the
    cache stays bounded (least recently used graphs are destroyed), and every
    step matches the oracle's."""
    say("graphs over changing time steps 128x64")
    cfg = PhysicalConfig(m=128, n=64)
    p, gpu, orc = tp.make_pair(vdi, oracle, cfg, tol=1e-12)
    dts = [p.dt_reg * f for f in (1.0, 0.9, 0.8, 0.7, 0.6, 1.0, 0.5, 0.9, 0.45, 0.4, 0.8, 0.35, 1.0, 0.3)]
    worst = 0.0
    for k, dt in enumerate(dts + dts):
        gpu.step_forward(dt)
        orc.step_forward(dt)
        worst = max(worst, tp.compare(gpu, orc, f"step {k} dt {dt:.3e}"))
    g, r = gpu.counters()
    say("  captured", g, "redone", r, "worst", worst)
    assert g >= len(dts)
    gpu.close()
    orc.close()


def test_gpu_rejected_schedule_leaves_settings(vdi):
    """A rejected schedule leaves every solver's sweeps as they were
    (per solver and for all four at once)."""
    p = PhysicalConfig(m=128, n=64).params()
    a = vdi.Fluid2D(p)
    before = [a.schedule(k) for k in (vdi.MAC, vdi.FEM, vdi.VX, vdi.VY)]
    for bad in [(5, 2), (3, 2, 5, 2), (0, 0)]:
        with pytest.raises(vdi.VdiError):
            a.set_schedule(vdi.FEM, *bad)
    with pytest.raises(vdi.VdiError):
        a.set_solver(nu1=5, nu2=2)
    assert [a.schedule(k) for k in (vdi.MAC, vdi.FEM, vdi.VX, vdi.VY)] == before
    a.close()


def test_gpu_tracers_match_oracle(vdi, oracle):
    """fluid_2d::update_tracers (fluid_2d.cc:507-525) on the GPU, moved at the
    start of every step (direct and captured), against the oracle's tracers
    over ten steps; tracers include ones outside the cell centres."""
    say("tracers 64x32")
    cfg = PhysicalConfig(m=64, n=32)
    p, gpu, orc = tp.make_pair(vdi, oracle, cfg, tol=1e-13)
    rng = np.random.default_rng(17)
    pos = np.column_stack([rng.uniform(p.ax - 0.02 * (p.bx - p.ax), p.bx, 512),
                           rng.uniform(p.ay, p.by * 1.01, 512)])
    gpu.set_tracers(pos)
    orc.set_tracers(pos)
    for k in range(10):
        gpu.step_forward(p.dt_reg)
        orc.step_forward(p.dt_reg)
        tp.compare(gpu, orc, f"step {k}")
        d = gpu.tracers() - orc.tracers()
        moved = np.abs(orc.tracers() - pos).max()
        assert np.abs(d).max() <= 1e-9 * moved + 1e-18, (k, np.abs(d).max(), moved)
    assert gpu.counters()[0] > 0
    gpu.close()
    orc.close()
