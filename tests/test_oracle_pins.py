"""Pins the CPU FP64 oracle against facts that do not come from its own code:
closed-form eigenpairs, exact integrals, the reference's tri_test known
answer, scipy's LAPACK dgtsv, Riemann-problem solutions, the PDE the gas layer
discretises, and exactness properties of the limiter and ENO2 stencils.

The reference itself cannot be compiled here (its multigrid library TGMG and
LAPACK are absent), so these checks are what stands behind every oracle
function; the whole coupled step is only pinned through its parts.
"""
import numpy as np
import pytest

import closed_forms as cf
from vdropimpact_amd.params import PhysicalConfig, drop_impact_params, h0_nd_for_hmin


# ---------------------------------------------------------------- operators

@pytest.mark.parametrize("m,n,dx,dy", [(16, 8, 0.1, 0.07), (24, 12, 1.0, 1.0)])
def test_mac_operator_eigenvectors(oracle, m, n, dx, dy):
    xxsp, yysp = 1 / dx ** 2, 1 / dy ** 2
    g = oracle.OracleMG(cf.MAC, m, n, xxsp, yysp)
    for k, l in [(0, 0), (1, 0), (3, 2), (m - 1, n - 1), (5, 7)]:
        z, az = cf.mac_mode(m, n, k, l, xxsp, yysp)
        np.testing.assert_allclose(g.apply(z), az, rtol=0, atol=1e-11 * np.abs(az).max())


@pytest.mark.parametrize("x_sym", [1, 0])
@pytest.mark.parametrize("kind", [cf.VX, cf.VY])
def test_visco_operator_eigenvectors(oracle, kind, x_sym):
    m, n, vfx, vfy = 18, 10, 0.3, 0.17
    g = oracle.OracleMG(kind, m, n, vfx, vfy, x_sym)
    for k, l in [(1, 0), (2, 3), (m - 1, n - 1), (7, 4)]:
        z, az = cf.visco_mode(m, n, k, l, vfx, vfy, x_sym, kind == cf.VX)
        np.testing.assert_allclose(g.apply(z), az, rtol=0, atol=1e-13 * np.abs(az).max())


@pytest.mark.parametrize("M,N,dx,dy", [(17, 9, 0.1, 0.07), (13, 13, 1.0, 1.0), (9, 17, 0.05, 0.2)])
def test_fem_operator_eigenvectors(oracle, M, N, dx, dy):
    g = oracle.OracleMG(cf.FEM, M, N, dx, dy)
    for k, l in [(0, 0), (1, 0), (3, 2), (M - 1, N - 2), (4, 1)]:
        z, az = cf.fem_mode(M, N, k, l, dx, dy)
        np.testing.assert_allclose(g.apply(z), az, rtol=0, atol=1e-12 * max(1.0, np.abs(az).max()))


def test_fem_dirichlet_row_is_identity(oracle):
    M, N = 9, 5
    g = oracle.OracleMG(cf.FEM, M, N, 0.1, 0.1)
    z = np.zeros(M * N)
    z[:M] = np.arange(M) + 1.0
    az = g.apply(z).reshape(N, M)
    fm = 4.0 / 3.0 * 2
    acc = np.where((np.arange(M) > 0) & (np.arange(M) < M - 1), 1.0, 0.5) * 0.5 * fm
    np.testing.assert_allclose(az[0], acc * z[:M], rtol=1e-15)


# ---------------------------------------------------------------- multigrid

@pytest.mark.parametrize("kind,m,n,cx,cy", [
    (cf.MAC, 64, 32, 1 / 0.01 ** 2, 1 / 0.013 ** 2),
    (cf.FEM, 65, 33, 0.01, 0.013),
    (cf.VX, 64, 32, 0.13, 0.09),
    (cf.VY, 64, 32, 0.13, 0.09),
])
def test_mg_solves_closed_form(oracle, kind, m, n, cx, cy):
    g = oracle.OracleMG(kind, m, n, cx, cy, 1)
    if kind == cf.MAC:
        z, b = cf.mac_mode(m, n, 3, 2, cx, cy)
    elif kind == cf.FEM:
        z, b = cf.fem_mode(m, n, 3, 2, cx, cy)
    else:
        z, b = cf.visco_mode(m, n, 3, 2, cx, cy, 1, kind == cf.VX)
    sol, cycles = g.solve(b, tol=1e-13)
    assert cycles > 0
    assert np.linalg.norm(sol - z) / np.linalg.norm(z) < 1e-10


@pytest.mark.parametrize("x_sym", [1, 0])
@pytest.mark.parametrize("wide", [True, False])
@pytest.mark.parametrize("kind", [cf.MAC, cf.FEM, cf.VX, cf.VY])
def test_mg_matches_dense_solve(oracle, kind, wide, x_sym):
    """A level of at most 3072 unknowns is the coarsest one and is solved
    exactly (separation of variables, transformed along the shorter side:
    rows for the wide grid, columns for the tall one): one V-cycle from zero
    gives numpy's dense solve of the assembled operator."""
    m, n = (16, 8) if wide else (10, 22)
    if kind == cf.FEM:
        m, n = m + 1, n + 1
    cx, cy = (0.1, 0.07) if kind == cf.FEM else ((100.0, 150.0) if kind == cf.MAC else (0.4, 0.3))
    g = oracle.OracleMG(kind, m, n, cx, cy, x_sym)
    assert g.levels == 1
    N = m * n
    A = np.stack([g.apply(np.eye(N)[c]) for c in range(N)], axis=1)
    rng = np.random.default_rng(7)
    b = rng.standard_normal(N)
    z_dense = np.linalg.solve(A, b)
    z, cycles = g.solve(b, tol=1e-12)
    assert cycles == 1
    np.testing.assert_allclose(z, z_dense, rtol=0, atol=1e-11 * np.abs(z_dense).max())


def test_mg_convergence_rate(oracle):
    """Each V-cycle reduces the residual by a factor of at least 5."""
    g = oracle.OracleMG(cf.MAC, 128, 64, 1.0, 1.0)
    b = np.random.default_rng(3).standard_normal(128 * 64)
    z = np.zeros_like(b)
    r0 = g.resnorm(b, z)
    z = g.vcycles(b, z, 4)
    assert g.resnorm(b, z) < r0 * 5.0 ** -4


# ---------------------------------------------------------------- tridiagonal

def test_tri_known_answer_thomas_and_pivot(oracle):
    """tri_test.cc: u''=1 on [0,1], u(0)=u(1)=0, 32 points; second-order
    differences are exact for quadratics, so u_i = x_i (x_i - 1)/2."""
    m = 32
    dx = 1.0 / (m + 1)
    xxsp = 1 / dx ** 2
    x = (np.arange(m) + 1) * dx
    exact = x * (x - 1) / 2
    a = np.full(m, xxsp); b = np.full(m, -2 * xxsp); c = np.full(m, xxsp); d = np.ones(m)
    L = oracle.lib()
    L.vdo_tri_solve_thomas(m, oracle.ptr(a), oracle.ptr(b), oracle.ptr(c), oracle.ptr(d))
    np.testing.assert_allclose(d, exact, rtol=1e-12, atol=1e-14)
    dl = np.full(m - 1, xxsp); dd = np.full(m, -2 * xxsp); du = np.full(m - 1, xxsp); rhs = np.ones(m)
    assert L.vdo_tri_solve_pivot(m, oracle.ptr(dl), oracle.ptr(dd), oracle.ptr(du), oracle.ptr(rhs)) == 0
    np.testing.assert_allclose(rhs, exact, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dominant", [True, False])
def test_tri_pivot_matches_lapack_dgtsv(oracle, dominant):
    from scipy.linalg.lapack import dgtsv
    rng = np.random.default_rng(11 + dominant)
    n = 200
    dl = rng.standard_normal(n - 1); du = rng.standard_normal(n - 1)
    d = rng.standard_normal(n) + (6.0 if dominant else 0.0)
    b = rng.standard_normal(n)
    ref_dl, ref_d, ref_du, ref_x, info = dgtsv(dl, d, du, b)
    assert info == 0
    a1, a2, a3, a4 = dl.copy(), d.copy(), du.copy(), b.copy()
    assert oracle.lib().vdo_tri_solve_pivot(n, oracle.ptr(a1), oracle.ptr(a2), oracle.ptr(a3), oracle.ptr(a4)) == 0
    np.testing.assert_allclose(a4, ref_x, rtol=1e-10, atol=1e-12 * np.abs(ref_x).max())


def test_tri_pivot_reports_singular(oracle):
    n = 4
    dl = np.zeros(n - 1); d = np.zeros(n); du = np.zeros(n - 1); b = np.ones(n)
    assert oracle.lib().vdo_tri_solve_pivot(n, oracle.ptr(dl), oracle.ptr(d), oracle.ptr(du), oracle.ptr(b)) != 0


# ---------------------------------------------------------------- limiter / Godunov

def test_mono_diff_linear_exact(oracle):
    md = oracle.lib().vdo_mono_diff
    for a, b in [(0.0, 1.0), (3.0, -0.25), (-1.0, 2.5)]:
        f = [a + b * k for k in range(5)]
        assert md(*f) == pytest.approx(b, rel=1e-14)


def test_mono_diff_extrema_and_symmetry(oracle):
    md = oracle.lib().vdo_mono_diff
    assert md(0, 1, 2, 1, 0) == 0.0
    assert md(0, -1, -2, -1, 0) == 0.0
    assert md(0, 0, 0, 1, 1) == 0.0           # limited at the foot of a step
    rng = np.random.default_rng(5)
    for _ in range(200):
        f = rng.standard_normal(5)
        s = md(*f)
        assert md(*(-f)) == pytest.approx(-s, abs=1e-15)
        assert md(*f[::-1]) == pytest.approx(-s, abs=1e-15)
        assert md(*(f + 3.7)) == pytest.approx(s, abs=1e-12)


def test_mono_diff_fourth_order(oracle):
    """On smooth monotone data the slope is the fourth-order central
    difference: error in h*f'(x) is O(h^5)."""
    md = oracle.lib().vdo_mono_diff
    x0 = 0.3
    errs = []
    for h in [0.02, 0.01, 0.005]:
        f = [np.sin(x0 + (k - 2) * h) for k in range(5)]
        errs.append(abs(md(*f) - h * np.cos(x0)))
    rates = [np.log2(errs[k] / errs[k + 1]) for k in range(2)]
    assert min(rates) > 4.7


def _gset(oracle, fn, u0, u1, v0, v1):
    vals = [np.array([x], np.float64) for x in (u0, u1, v0, v1)]
    getattr(oracle.lib(), fn)(*[oracle.ptr(v) for v in vals])
    return [float(v[0]) for v in vals]


def test_godunov_set_is_burgers_riemann(oracle):
    rng = np.random.default_rng(2)
    pairs = [(1, 0.5), (-0.5, -1), (-1, 1), (1, -1), (2, -1), (1, -2), (0, 0), (0.5, 2), (-2, -0.5)]
    pairs += [tuple(x) for x in rng.standard_normal((200, 2))]
    for u0, u1 in pairs:
        v0, v1 = 0.3, -0.7
        val, side = cf.burgers_godunov(u0, u1)
        r = _gset(oracle, "vdo_godunov_set", u0, u1, v0, v1)
        assert r[0] == u0 and r[2] == v0
        assert r[1] == val
        assert r[3] == (v0 if side == 0 else v1 if side == 1 else 0.5 * (v0 + v1))


def test_godunov_set_tang(oracle):
    rng = np.random.default_rng(4)
    for u0, u1 in [tuple(x) for x in rng.standard_normal((200, 2))] + [(-1, 1), (1, -1)]:
        v0, v1 = 0.25, 0.75
        val, side = cf.burgers_godunov(u0, u1)
        r = _gset(oracle, "vdo_godunov_set_tang", u0, u1, v0, v1)
        assert r[1] == val
        # the left value becomes the upwind state, or the average at a sonic point
        assert r[0] == (u0 if side == 0 else u1 if side == 1 else 0.5 * (u0 + u1))
        assert r[3] == (v0 if side == 0 else v1 if side == 1 else 0.5 * (v0 + v1))


# ---------------------------------------------------------------- Simpson BC

@pytest.mark.parametrize("x_sym", [1, 0])
def test_simpson_constant_pressure_converges_4th_order(oracle, x_sym):
    ax, bx, p0 = (0.0 if x_sym else -1.0), 1.0, 2.5
    L = oracle.lib()
    for (x, y) in [(1.05, 0.3), (0.4, 1.2)]:
        du_ex, dv_ex = cf.simpson_exact(x, y, ax, bx, x_sym, p0)
        errs = []
        for m in [32, 64, 128]:
            me = m + 1
            dx = (bx - ax) / m
            p = np.full(me, p0)
            du = np.zeros(1); dv = np.zeros(1)
            L.vdo_simpson_velocity(me, ax, dx, x_sym, oracle.ptr(p), x, y, dx / 3, oracle.ptr(du), oracle.ptr(dv))
            errs.append((abs(du[0] - du_ex), abs(dv[0] - dv_ex)))
        for c in range(2):
            assert errs[-1][c] < 1e-6 * max(abs(du_ex), abs(dv_ex))
            assert np.log2(errs[0][c] / errs[1][c]) > 3.5


# ---------------------------------------------------------------- gas layer

def _gas_sim(oracle, m=64, n=32, x_sym=0):
    p = PhysicalConfig(m=m, n=n, x_sym=bool(x_sym)).params()
    s = oracle.OracleSim(p)
    s.init_fields()
    return s, p


def _height_setup(oracle, p, h, U, vc):
    m, n = p.m, p.n
    u = np.zeros((n + 4, m + 4)); v = np.zeros((n + 4, m + 4))
    u[1:3, 2:m + 2] = U          # rows -1 and 0
    v[1:3, 2:m + 2] = vc
    s = oracle.OracleSim(p)
    s.init_fields()
    s.set(oracle.H, h); s.set(oracle.U, u); s.set(oracle.V, v)
    return s


@pytest.mark.parametrize("U", [0.3, -0.3])
def test_eno2_height_update_exact_for_linear_h(oracle, U):
    """ENO2 is exact for linear h, so h_t = -U h' + v - V (with a time step
    small enough that the in-place sweep does not perturb the stencil)."""
    p = PhysicalConfig(m=64, n=32, x_sym=False).params()
    dx = (p.bx - p.ax) / p.m
    x = p.ax + (np.arange(p.m) + 0.5) * dx
    a, bslope = 4e-6, 0.01
    s = _height_setup(oracle, p, a + bslope * x, U, 0.1)
    fbd0 = s.get(oracle.FBD)
    dt = 1e-16
    s.gas_height(dt)
    ht_exact = -U * bslope + 0.1 - p.V
    np.testing.assert_allclose(s.get(oracle.HT), ht_exact, rtol=1e-8)
    np.testing.assert_allclose(s.get(oracle.FBD), fbd0 + dt * 0.1, rtol=1e-12)
    np.testing.assert_allclose(s.get(oracle.HX), bslope, rtol=1e-6)


def test_height_update_sweeps_in_place(oracle):
    """The reference updates h inside the upward sweep (gas_layer_solve.cc:
    152-159), so for constant h and U>0 the second point sees the already
    advanced h[0]: with c = v - V, d = dt c and uc = U/(2 dx), ENO2 picks the
    one-sided branch 3h - 4(h+d) + (h+2d) = -2d and h_t[1] = c + 2 uc d."""
    p = PhysicalConfig(m=64, n=32, x_sym=False).params()
    dx = (p.bx - p.ax) / p.m
    U, vc, dt = 0.3, 0.1, 1e-7
    s = _height_setup(oracle, p, np.full(p.m, 4e-6), U, vc)
    s.gas_height(dt)
    ht = s.get(oracle.HT)
    c = vc - p.V
    assert ht[0] == pytest.approx(c, rel=1e-12)
    assert ht[1] == pytest.approx(c + 2 * (0.5 * U / dx) * dt * c, rel=1e-9)
    assert abs(ht[1] - c) > 1e-4 * abs(c)


def _lubrication_residual(p, s, dt):
    """F(pg) for the compressible lubrication equation
        d/dt(rho h) = d/dx(rho h^3/(12 mu_g) dP/dx),  rho ~ P^alpha,  P = pg + Pamb,
    multiplied by 12 mu_g P^(1-alpha) and written with centred differences:
        alpha h^3 P_x^2 + P (3 h^2 h_x P_x + h^3 P_xx - 12 mu_g h_t)
            - 12 mu_g alpha h (pg - pg_old)/dt
    at the interior nodes, with h, h_t averaged to the nodes."""
    m = p.m
    me = m + 1
    xsp = m / (p.bx - p.ax)
    h, ht, hx = s.get(14), s.get(16), s.get(15)
    pg, pgo = s.get(18), s.get(19)
    i = np.arange(1, me - 1)
    hn = 0.5 * (h[i - 1] + h[np.minimum(i, m - 1)])
    htn = 0.5 * (ht[i - 1] + ht[np.minimum(i, m - 1)])
    P = pg + p.Pamb
    Px = 0.5 * xsp * (P[i + 1] - P[i - 1])
    Pxx = xsp * xsp * (P[i + 1] - 2 * P[i] + P[i - 1])
    F = (p.alpha * hn ** 3 * Px ** 2 + P[i] * (3 * hn ** 2 * hx[i] * Px + hn ** 3 * Pxx - 12 * p.mug * htn)
         - 12 * p.mug * p.alpha * hn * (pg[i] - pgo[i]) / dt)
    scale = 12 * p.mug * p.alpha * hn * P[i] / dt
    return F, scale


def test_newton_root_satisfies_lubrication_equation(oracle):
    s, p = _gas_sim(oracle, m=128, n=64, x_sym=1)
    for _ in range(3):
        s.step_forward(p.dt_reg)
    s.gas_update(s.time, p.dt_reg)
    assert 2 <= s.last_newton() <= 20
    F, scale = _lubrication_residual(p, s, p.dt_reg)
    assert np.abs(F).max() < 1e-10 * np.abs(scale).max()
    pg = s.get(oracle.PG)
    assert pg[-1] == 0.0                         # Dirichlet at the far end
    assert np.abs(pg).max() > 1.0                # the gas is compressed


def test_liquid_pressure_and_tau_bc(oracle):
    s, p = _gas_sim(oracle, m=64, n=32, x_sym=1)
    s.step_forward(p.dt_reg)
    tau = s.get(oracle.GTAU)
    ml = p.m + 4
    # gas_layer.cc:25-30: odd reflections at both ends
    assert tau[0] == -tau[3] and tau[1] == -tau[2]
    assert tau[ml - 2] == -tau[ml - 3] and tau[ml - 1] == -tau[ml - 4]
    # the model never writes tau[m], tau[m+1] (consumed two cells shifted)
    assert tau[p.m] == 0.0 and tau[p.m + 1] == 0.0


# ---------------------------------------------------------------- configuration

def test_config_nondimensionalisation():
    """h0_nd=4.586010, L_nd=30 give h_min=5 um and L=0.9906 mm for the
    32 cSt, R=1 mm, V=0.5 m/s drop (st = mug/(rho V R) = 3.6e-5)."""
    p = drop_impact_params(256, 128)
    assert p.h0 == pytest.approx(5e-6, rel=1e-7)
    assert p.bx == pytest.approx(9.906e-4, rel=1e-4)
    assert p.by == pytest.approx(p.bx / 2)
    assert p.bx / 256 * 8e-3 == pytest.approx(p.dt_reg)
    assert p.dt_reg == pytest.approx(3.096e-8, rel=1e-3)
    assert p.mu == pytest.approx(0.032)
    assert h0_nd_for_hmin(5e-6, 1.8e-5, 1000, 0.5, 1e-3) == pytest.approx(4.586010, rel=1e-6)
    q = PhysicalConfig(m=64, n=32, x_sym=False, implicit_visc=False).params()
    assert q.ax == -q.bx and q.by == pytest.approx(q.bx)   # full width doubles by
    dx = (q.bx - q.ax) / 64
    assert q.dt_reg == pytest.approx(dx * dx * q.rho / q.mu * 8e-3 * 0.25)


# ---------------------------------------------------------------- whole step

def test_oracle_step_runs_and_conserves_symmetry(oracle):
    """x_sym run: left ghosts are the mirror images of the first columns."""
    s, p = _gas_sim(oracle, m=64, n=32, x_sym=1)
    for _ in range(5):
        s.step_forward(p.dt_reg)
    u = s.padded(oracle.U); v = s.padded(oracle.V)
    np.testing.assert_array_equal(u[2:-2, 1], -u[2:-2, 2])
    np.testing.assert_array_equal(v[2:-2, 1], v[2:-2, 2])
    for k in range(4):
        assert 0 <= s.last_cycles(k) < 30
    assert np.isfinite(u).all() and np.abs(v).max() > 0
