"""Pins of the oracle's step, pass by pass, against facts that do not come
from the oracle's own formulas: exactness on fields the discretisation must
reproduce (linear and bilinear data, uniform flow), discrete identities
(a discrete curl has zero discrete divergence; after the MAC correction the
face divergence is the MAC residual), the boundary conditions each ghost
value stands for (stress, incompressibility, mirror symmetry, Dirichlet and
Neumann data of the Helmholtz solves), closed-form integrals, and second-
order convergence of the predictor to the Taylor-Green vortex, an exact
solution of the incompressible Navier-Stokes equations.

Each pass is run alone through vdo_pass (oracle/vdo.h) on arrays the test
sets, following fluid_2d.cc:93-405, fluid_2d_sio.cc:256-343 and
gas_layer_solve.cc:32-249 / gas_layer.cc:25-30 / mgs_common.cc:17-50.
"""
import numpy as np
import pytest

from vdropimpact_amd.params import PhysicalConfig, VdiParams


def mkparams(m, n, ax=0.0, bx=1.0, ay=0.0, by=1.0, rho=1.0, mu=0.1, x_sym=0, implicit=0, dt=1e-3, **gas):
    g = dict(alpha=1 / 1.4, mug=1.8e-5, Pamb=1e5, h0=5e-6, R=1e-3, V=0.5, sigma=0.02)
    g.update(gas)
    return VdiParams(m=m, n=n, x_sym=x_sym, implicit_visc=implicit, ax=ax, bx=bx, ay=ay, by=by, rho=rho, mu=mu,
                     dt_reg=dt, **g)


class Grid:
    """Padded (n+4) x (m+4) arrays: cell (i, j), i in [-2, m+1], j in
    [-2, n+1], at [j+2, i+2]; cell centres and (for the nodal pressure kept
    in the same layout) node positions."""

    def __init__(self, p):
        self.m, self.n = p.m, p.n
        self.ax, self.ay = p.ax, p.ay
        self.dx = (p.bx - p.ax) / p.m
        self.dy = (p.by - p.ay) / p.n
        i = np.arange(-2, p.m + 2)
        j = np.arange(-2, p.n + 2)
        self.XC, self.YC = np.meshgrid(p.ax + (i + 0.5) * self.dx, p.ay + (j + 0.5) * self.dy)
        self.XN, self.YN = np.meshgrid(p.ax + i * self.dx, p.ay + j * self.dy)

    def at(self, a, i, j):
        return a[j + 2, i + 2]

    def inner(self, a):
        return a[2:self.n + 2, 2:self.m + 2]


def sim(oracle, p):
    return oracle.OracleSim(p, rel_tol=1e-13)


def pad(s, aid):
    return s.padded(aid).copy()


# ====================================================================== K1

def test_k1_linear_fields_extrapolate_along_the_normal_characteristic(oracle):
    """fluid_2d.cc:105-132.  For linear u, v the limited slopes are the exact
    gradients, and each face value is the field traced back from the face
    along the face-normal velocity over half a step: u_left = u(x - dx/2 -
    dt/2 u, y), u_down = u(x, y - dy/2 - dt/2 v), and so on."""
    p = mkparams(16, 12, bx=1.0, by=0.9)
    G, s = Grid(p), sim(oracle, p)
    a, b = (0.7, 0.3, -0.4), (-0.2, 0.5, 0.25)

    def U(x, y):
        return a[0] + a[1] * x + a[2] * y

    def W(x, y):
        return b[0] + b[1] * x + b[2] * y
    uc, vc = U(G.XC, G.YC), W(G.XC, G.YC)
    s.set(oracle.U, uc)
    s.set(oracle.V, vc)
    dt = 0.013
    s.run_pass(oracle.PASS_K1, dt)
    for aid, val in ((oracle.UX, a[1]), (oracle.UY, a[2]), (oracle.VXS, b[1]), (oracle.VYS, b[2])):
        np.testing.assert_allclose(G.inner(pad(s, aid)), val, rtol=1e-12)
    X, Y, dx, dy = G.XC, G.YC, G.dx, G.dy
    expect = {
        oracle.UL: U(X - dx / 2 - dt / 2 * uc, Y), oracle.UR: U(X + dx / 2 - dt / 2 * uc, Y),
        oracle.VL: W(X - dx / 2 - dt / 2 * uc, Y), oracle.VR: W(X + dx / 2 - dt / 2 * uc, Y),
        oracle.UD: U(X, Y - dy / 2 - dt / 2 * vc), oracle.UU: U(X, Y + dy / 2 - dt / 2 * vc),
        oracle.VD: W(X, Y - dy / 2 - dt / 2 * vc), oracle.VU: W(X, Y + dy / 2 - dt / 2 * vc),
    }
    for aid, e in expect.items():
        np.testing.assert_allclose(G.inner(pad(s, aid)), G.inner(e), rtol=1e-13, atol=1e-14)


def test_k1_slopes_use_their_own_direction(oracle):
    """The x slope of u comes from u's x neighbours and vanishes at an x
    extremum even where u varies in y (and the y slope of v likewise)."""
    p = mkparams(16, 12, bx=1.0, by=0.9)
    G, s = Grid(p), sim(oracle, p)
    ic, jc = 5, 4
    xc, yc = G.at(G.XC, ic, 0), G.at(G.YC, 0, jc)
    s.set(oracle.U, (G.XC - xc) ** 2 + 0.8 * G.YC)
    s.set(oracle.V, (G.YC - yc) ** 2 - 0.6 * G.XC)
    s.run_pass(oracle.PASS_K1, 0.01)
    ux, uy, vx, vy = (pad(s, k) for k in (oracle.UX, oracle.UY, oracle.VXS, oracle.VYS))
    assert np.all(ux[2:-2, ic + 2] == 0)
    assert np.all(vy[jc + 2, 2:-2] == 0)
    np.testing.assert_allclose(G.inner(uy), 0.8, rtol=1e-12)
    np.testing.assert_allclose(G.inner(vx), -0.6, rtol=1e-12)


# ====================================================================== K2/K5

def _faces(s, oracle):
    return {k: pad(s, k) for k in (oracle.UL, oracle.UR, oracle.VL, oracle.VR, oracle.UD, oracle.UU,
                                   oracle.VD, oracle.VU)}


@pytest.mark.parametrize("x_sym", [0, 1])
def test_k2_takes_the_upwind_states_rightward_upward(oracle, x_sym):
    """fluid_2d.cc:136-150.  Where the flow crosses every face to the right
    and upwards, each face carries its upwind (left, lower) cell's state; the
    outflow faces at the right and top take the interior state; the left and
    bottom faces keep the first cell's own state (zero normal velocity at the
    symmetry line with x_sym)."""
    p = mkparams(16, 12, bx=1.0, by=0.9, x_sym=x_sym)
    G, s = Grid(p), sim(oracle, p)
    s.set(oracle.U, 1.0 + 0.2 * G.XC + 0.1 * G.YC)
    s.set(oracle.V, 0.8 + 0.1 * G.XC - 0.2 * G.YC)
    s.run_pass(oracle.PASS_K1, 0.01)
    f1 = _faces(s, oracle)
    s.run_pass(oracle.PASS_K2)
    f2 = _faces(s, oracle)
    m, n = p.m, p.n
    R = slice(2, n + 2)     # rows 0..n-1
    Cc = slice(2, m + 2)    # columns 0..m-1
    # vertical faces i+1/2, i = 0..m-2: right cell's left state := left cell's right state
    np.testing.assert_array_equal(f2[oracle.UL][R, 3:m + 2], f1[oracle.UR][R, 2:m + 1])
    np.testing.assert_array_equal(f2[oracle.VL][R, 3:m + 2], f1[oracle.VR][R, 2:m + 1])
    np.testing.assert_array_equal(f2[oracle.UR][R, 2:m + 2], f1[oracle.UR][R, 2:m + 2])
    # right outflow face m-1/2
    np.testing.assert_array_equal(f2[oracle.UL][R, m + 2], f1[oracle.UR][R, m + 1])
    np.testing.assert_array_equal(f2[oracle.VL][R, m + 2], f1[oracle.VR][R, m + 1])
    # left face -1/2
    left = np.zeros(n) if x_sym else f1[oracle.UL][R, 2]
    np.testing.assert_array_equal(f2[oracle.UL][R, 2], left)
    np.testing.assert_array_equal(f2[oracle.UR][R, 1], left)
    # horizontal faces j+1/2
    np.testing.assert_array_equal(f2[oracle.VD][3:n + 2, Cc], f1[oracle.VU][2:n + 1, Cc])
    np.testing.assert_array_equal(f2[oracle.UD][3:n + 2, Cc], f1[oracle.UU][2:n + 1, Cc])
    np.testing.assert_array_equal(f2[oracle.VU][2:n + 2, Cc], f1[oracle.VU][2:n + 2, Cc])
    np.testing.assert_array_equal(f2[oracle.VD][n + 2, Cc], f1[oracle.VU][n + 1, Cc])
    np.testing.assert_array_equal(f2[oracle.UD][n + 2, Cc], f1[oracle.UU][n + 1, Cc])
    np.testing.assert_array_equal(f2[oracle.VU][1, Cc], f1[oracle.VD][2, Cc])


def test_k2_takes_the_upwind_states_leftward_downward(oracle):
    """Flow to the left and downwards: each face's state comes from the right
    (upper) cell; the tangential-term form records it in the left cell's
    right state (fluid_2d.cc:469-476)."""
    p = mkparams(16, 12, bx=1.0, by=0.9)
    G, s = Grid(p), sim(oracle, p)
    s.set(oracle.U, -1.0 - 0.2 * G.XC + 0.1 * G.YC)
    s.set(oracle.V, -0.8 + 0.1 * G.XC - 0.2 * G.YC)
    s.run_pass(oracle.PASS_K1, 0.01)
    f1 = _faces(s, oracle)
    s.run_pass(oracle.PASS_K2)
    f2 = _faces(s, oracle)
    m, n = p.m, p.n
    R, Cc = slice(2, n + 2), slice(2, m + 2)
    np.testing.assert_array_equal(f2[oracle.UR][R, 2:m + 1], f1[oracle.UL][R, 3:m + 2])
    np.testing.assert_array_equal(f2[oracle.UL][R, 3:m + 2], f1[oracle.UL][R, 3:m + 2])
    np.testing.assert_array_equal(f2[oracle.VL][R, 3:m + 2], f1[oracle.VL][R, 3:m + 2])
    np.testing.assert_array_equal(f2[oracle.VU][2:n + 1, Cc], f1[oracle.VD][3:n + 2, Cc])
    np.testing.assert_array_equal(f2[oracle.VD][3:n + 2, Cc], f1[oracle.VD][3:n + 2, Cc])
    np.testing.assert_array_equal(f2[oracle.UD][3:n + 2, Cc], f1[oracle.UD][3:n + 2, Cc])


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_k5_takes_the_upwind_states(oracle, sign):
    """fluid_2d.cc:205-215: the same upwind choice on the full predictor's
    states, without the tangential bookkeeping (the downwind side's own
    states are never changed), and the same outflow copies."""
    p = mkparams(12, 10, bx=1.0, by=0.9)
    G, s = Grid(p), sim(oracle, p)
    rng = np.random.default_rng(3)
    shape = G.XC.shape
    vals = {k: sign * (0.5 + rng.random(shape)) for k in (oracle.UL, oracle.UR, oracle.VD, oracle.VU)}
    vals.update({k: rng.standard_normal(shape) for k in (oracle.VL, oracle.VR, oracle.UD, oracle.UU)})
    for k, v in vals.items():
        s.set(k, v)
    s.run_pass(oracle.PASS_K5)
    f = _faces(s, oracle)
    m, n = p.m, p.n
    R, Cc = slice(2, n + 2), slice(2, m + 2)
    for a in (oracle.UR, oracle.VR, oracle.VU, oracle.UU):
        np.testing.assert_array_equal(f[a], vals[a])       # upwind-side states untouched
    if sign > 0:
        np.testing.assert_array_equal(f[oracle.UL][R, 3:m + 2], vals[oracle.UR][R, 2:m + 1])
        np.testing.assert_array_equal(f[oracle.VL][R, 3:m + 2], vals[oracle.VR][R, 2:m + 1])
        np.testing.assert_array_equal(f[oracle.VD][3:n + 2, Cc], vals[oracle.VU][2:n + 1, Cc])
        np.testing.assert_array_equal(f[oracle.UD][3:n + 2, Cc], vals[oracle.UU][2:n + 1, Cc])
    else:
        for a in (oracle.UL, oracle.VL):
            np.testing.assert_array_equal(f[a][R, 3:m + 2], vals[a][R, 3:m + 2])
        for a in (oracle.VD, oracle.UD):
            np.testing.assert_array_equal(f[a][3:n + 2, Cc], vals[a][3:n + 2, Cc])
    np.testing.assert_array_equal(f[oracle.UL][R, m + 2], vals[oracle.UR][R, m + 1])
    np.testing.assert_array_equal(f[oracle.VL][R, m + 2], vals[oracle.VR][R, m + 1])
    np.testing.assert_array_equal(f[oracle.VD][n + 2, Cc], vals[oracle.VU][n + 1, Cc])
    np.testing.assert_array_equal(f[oracle.UD][n + 2, Cc], vals[oracle.UU][n + 1, Cc])


# ====================================================================== K3

def test_k3_is_the_upwind_flux_difference(oracle):
    """fluid_2d.cc:153-160 on linear rightward/upward flow after K1, K2:
    c0 = d(u^2/2)/dx and c3 = d(v^2/2)/dy as differences of the upwind face
    states, c1 = u dv/dx and c2 = v du/dy with the face-averaged transport
    velocity."""
    p = mkparams(16, 12, bx=1.0, by=0.9)
    G, s = Grid(p), sim(oracle, p)
    s.set(oracle.U, 1.0 + 0.2 * G.XC + 0.1 * G.YC)
    s.set(oracle.V, 0.8 + 0.1 * G.XC - 0.2 * G.YC)
    s.run_pass(oracle.PASS_K1, 0.01)
    f1 = _faces(s, oracle)
    s.run_pass(oracle.PASS_K2)
    s.run_pass(oracle.PASS_K3)
    m, n, dx, dy = p.m, p.n, G.dx, G.dy
    R, Cc = slice(2, n + 2), slice(2, m + 2)
    # upwind states on the vertical faces -1/2 .. m-1/2 (columns) and the
    # horizontal faces -1/2 .. n-1/2 (rows)
    Fu = np.concatenate([f1[oracle.UL][R, 2:3], f1[oracle.UR][R, 2:m + 2]], axis=1)
    Fv = np.concatenate([f1[oracle.VL][R, 2:3], f1[oracle.VR][R, 2:m + 2]], axis=1)
    Hv = np.concatenate([f1[oracle.VD][2:3, Cc], f1[oracle.VU][2:n + 2, Cc]], axis=0)
    Hu = np.concatenate([f1[oracle.UD][2:3, Cc], f1[oracle.UU][2:n + 2, Cc]], axis=0)
    c0 = (Fu[:, 1:] ** 2 - Fu[:, :-1] ** 2) / (2 * dx)
    c1 = 0.5 * (Fu[:, 1:] + Fu[:, :-1]) * (Fv[:, 1:] - Fv[:, :-1]) / dx
    c2 = 0.5 * (Hv[1:] + Hv[:-1]) * (Hu[1:] - Hu[:-1]) / dy
    c3 = (Hv[1:] ** 2 - Hv[:-1] ** 2) / (2 * dy)
    for aid, e in ((oracle.C0, c0), (oracle.C1, c1), (oracle.C2, c2), (oracle.C3, c3)):
        np.testing.assert_allclose(G.inner(pad(s, aid)), e, rtol=1e-12, atol=1e-13)


# ====================================================================== Taylor-Green

def tg(x, y, t, nu, rho):
    """Taylor-Green vortex: u = sin x cos y F, v = -cos x sin y F,
    p = rho/4 (cos 2x + cos 2y) F^2, F = exp(-2 nu t); an exact solution of
    the incompressible Navier-Stokes equations."""
    F = np.exp(-2 * nu * t)
    return (np.sin(x) * np.cos(y) * F, -np.cos(x) * np.sin(y) * F,
            0.25 * rho * (np.cos(2 * x) + np.cos(2 * y)) * F * F)


def tg_setup(oracle, m, nu=0.1, rho=1.0, t=0.3, cfl=0.4):
    L = 2 * np.pi
    p = mkparams(m, m, bx=L, by=L, rho=rho, mu=nu * rho, x_sym=0, implicit=0)
    G, s = Grid(p), sim(oracle, p)
    u, v, _ = tg(G.XC, G.YC, t, nu, rho)
    _, _, pn = tg(G.XN, G.YN, t, nu, rho)
    s.set(oracle.U, u)
    s.set(oracle.V, v)
    s.set(oracle.P, pn)
    return p, G, s, cfl * G.dx


def _interior(a, G, k=3):
    return a[2 + k:G.n + 2 - k, 2 + k:G.m + 2 - k]


def test_predictor_k1_to_k4_converges_to_taylor_green_at_second_order(oracle):
    """K1-K4 (fluid_2d.cc:105-201) predict every face value at the half step.
    On the Taylor-Green vortex the left/right/down/up states converge to the
    exact velocity at the face and t + dt/2 at second order (dt ~ dx), which
    a dropped or mis-signed term (pressure, viscosity, transverse, normal
    advection) would spoil."""
    nu, rho, t = 0.1, 1.0, 0.3
    errs = []
    for m in (32, 64, 128):
        p, G, s, dt = tg_setup(oracle, m, nu, rho, t)
        for ps in (oracle.PASS_K1, oracle.PASS_K2, oracle.PASS_K3, oracle.PASS_K4):
            s.run_pass(ps, dt)
        th = t + dt / 2
        X, Y, dx, dy = G.XC, G.YC, G.dx, G.dy
        e = 0.0
        for aid, comp, (ox, oy) in ((oracle.UL, 0, (-dx / 2, 0)), (oracle.UR, 0, (dx / 2, 0)),
                                    (oracle.UD, 0, (0, -dy / 2)), (oracle.UU, 0, (0, dy / 2)),
                                    (oracle.VL, 1, (-dx / 2, 0)), (oracle.VR, 1, (dx / 2, 0)),
                                    (oracle.VD, 1, (0, -dy / 2)), (oracle.VU, 1, (0, dy / 2))):
            ex = tg(X + ox, Y + oy, th, nu, rho)[comp]
            e = max(e, np.abs(_interior(pad(s, aid) - ex, G)).max())
        errs.append(e)
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert errs[-1] < 2e-3 and np.all(rates > 1.8), (errs, rates)


def test_transverse_terms_converge_to_the_advective_products(oracle):
    """K3's c0..c3 approximate u du/dx, u dv/dx, v du/dy and v dv/dy of the
    Taylor-Green flow at the cell centres."""
    nu, rho, t = 0.1, 1.0, 0.3
    errs = []
    for m in (32, 64, 128):
        p, G, s, dt = tg_setup(oracle, m, nu, rho, t)
        for ps in (oracle.PASS_K1, oracle.PASS_K2, oracle.PASS_K3):
            s.run_pass(ps, dt)
        X, Y = G.XC, G.YC
        F = np.exp(-2 * nu * t)
        u, v = np.sin(X) * np.cos(Y) * F, -np.cos(X) * np.sin(Y) * F
        ux, uy = np.cos(X) * np.cos(Y) * F, -np.sin(X) * np.sin(Y) * F
        vx, vy = np.sin(X) * np.sin(Y) * F, -np.cos(X) * np.cos(Y) * F
        e = 0.0
        for aid, ex in ((oracle.C0, u * ux), (oracle.C1, u * vx), (oracle.C2, v * uy), (oracle.C3, v * vy)):
            e = max(e, np.abs(_interior(pad(s, aid) - ex, G)).max())
        errs.append(e)
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert errs[-1] < 0.05 and np.all(rates > 0.9), (errs, rates)


def test_advection_chain_k1_to_k8_converges_to_taylor_green(oracle):
    """The whole explicit part of a step (K1-K8 with the MAC projection)
    gives c = u(t) + dt (-u.grad u + nu lap u), i.e. the exact velocity at
    t + dt plus dt grad p / rho at the half step, to O(dt^2) per step."""
    nu, rho, t = 0.05, 1.0, 0.3
    errs = []
    for m in (32, 64, 128):
        p, G, s, dt = tg_setup(oracle, m, nu, rho, t)
        for ps in (oracle.PASS_K1, oracle.PASS_K2, oracle.PASS_K3, oracle.PASS_K4, oracle.PASS_K5,
                   oracle.PASS_K6):
            s.run_pass(ps, dt)
        s.solve_v_cycle(oracle.KIND_MAC, t + dt / 2)
        s.run_pass(oracle.PASS_K7)
        s.run_pass(oracle.PASS_K8, dt)
        X, Y = G.XC, G.YC
        u1, v1, _ = tg(X, Y, t + dt, nu, rho)
        F2 = np.exp(-4 * nu * (t + dt / 2))
        px, py = -0.5 * rho * np.sin(2 * X) * F2, -0.5 * rho * np.sin(2 * Y) * F2
        e = max(np.abs(_interior(pad(s, oracle.C0) - (u1 + dt * px / rho), G)).max(),
                np.abs(_interior(pad(s, oracle.C1) - (v1 + dt * py / rho), G)).max())
        errs.append(e / dt)
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 0.9), (errs, rates)


# ====================================================================== K6, K7

def test_k6_discrete_curl_has_zero_divergence_and_bottom_dirichlet_data(oracle):
    """fluid_2d.cc:218-232.  Face velocities from a nodal stream function
    have zero face divergence; the bottom row carries the face Dirichlet value
    of the MAC variable z = dt p / (2 rho), with p the gas pressure averaged
    over the face's two nodes and the step's two time levels."""
    p = mkparams(16, 12, bx=1.0, by=0.9, rho=997.0)
    G, s = Grid(p), sim(oracle, p)
    rng = np.random.default_rng(7)
    psi = np.sin(2.1 * G.XN) * np.cos(1.3 * G.YN) + 0.3 * G.XN * G.YN ** 2
    # padded index [j+2, i+2] of node (i, j); faces of cell (i, j):
    # ul at x_i - dx/2 (nodes (i,j), (i,j+1)), vd at y_j - dy/2 (nodes (i,j), (i+1,j))
    ul = np.zeros_like(psi)
    vd = np.zeros_like(psi)
    ul[:-1, :] = (psi[1:, :] - psi[:-1, :]) / G.dy
    vd[:, :-1] = -(psi[:, 1:] - psi[:, :-1]) / G.dx
    s.set(oracle.UL, ul)
    s.set(oracle.VD, vd)
    gp, gpo = rng.standard_normal(p.m + 1) * 50, rng.standard_normal(p.m + 1) * 50
    s.set(oracle.GP, gp)
    s.set(oracle.GP_OLD, gpo)
    dt = 2e-4
    s.run_pass(oracle.PASS_K6, dt)
    src = s.get(oracle.SRC)
    m, n, mn = p.m, p.n, p.m * p.n
    scale = np.abs(ul).max() / G.dx
    assert np.abs(src[m:mn]).max() < 1e-12 * scale
    sbc = dt / (2 * p.rho) * 0.25 * (gpo[:-1] + gpo[1:] + gp[:-1] + gp[1:])
    np.testing.assert_allclose(src[mn:mn + m], sbc, rtol=1e-14)
    np.testing.assert_allclose(src[:m], -2 / G.dy ** 2 * sbc, rtol=1e-12, atol=1e-12 * scale)


def test_k6_linear_faces_give_constant_divergence(oracle):
    p = mkparams(16, 12, bx=1.0, by=0.9)
    G, s = Grid(p), sim(oracle, p)
    s.set(oracle.UL, 2.5 * (G.XC - G.dx / 2))
    s.set(oracle.VD, -0.75 * (G.YC - G.dy / 2))
    s.run_pass(oracle.PASS_K6, 1e-3)
    src = s.get(oracle.SRC)
    np.testing.assert_allclose(src[p.m:p.m * p.n], 2.5 - 0.75, rtol=1e-12)


@pytest.mark.parametrize("x_sym", [0, 1])
def test_k7_leaves_the_mac_residual_as_face_divergence(oracle, x_sym):
    """fluid_2d.cc:218-248 with the MAC operator (mgs_mac.hh:25-50): after
    the correction, every cell's face divergence equals that cell's residual
    b - A z of the MAC system (so a converged solve leaves divergence-free
    faces), including the bottom row with its Dirichlet face."""
    p = mkparams(24, 16, bx=1.0, by=0.7, rho=997.0, x_sym=x_sym)
    G, s = Grid(p), sim(oracle, p)
    rng = np.random.default_rng(11)
    s.set(oracle.UL, np.sin(3 * G.XC) * np.cos(2 * G.YC) + 0.1 * rng.standard_normal(G.XC.shape))
    s.set(oracle.VD, np.cos(2 * G.XC) * G.YC + 0.1 * rng.standard_normal(G.XC.shape))
    s.set(oracle.GP, rng.standard_normal(p.m + 1) * 30)
    s.set(oracle.GP_OLD, rng.standard_normal(p.m + 1) * 30)
    dt = 3e-4
    s.run_pass(oracle.PASS_K6, dt)
    src = s.get(oracle.SRC)[:p.m * p.n].copy()
    # a solve stopped early on purpose, so that the residual is not small
    orig_tol = 1e-13
    s.solve_v_cycle(oracle.KIND_MAC, 1.0)
    z = s.get(oracle.Z_MAC)
    z_pert = z + 1e-3 * np.abs(z).max() * rng.standard_normal(z.size)
    s.set(oracle.Z_MAC, z_pert)
    s.run_pass(oracle.PASS_K7)
    ul, vd = pad(s, oracle.UL), pad(s, oracle.VD)
    m, n = p.m, p.n
    div = ((ul[2:n + 2, 3:m + 3] - ul[2:n + 2, 2:m + 2]) / G.dx +
           (vd[3:n + 3, 2:m + 2] - vd[2:n + 2, 2:m + 2]) / G.dy).ravel()
    A = oracle.OracleMG(oracle.KIND_MAC, m, n, 1 / G.dx ** 2, 1 / G.dy ** 2, x_sym)
    resid = src - A.apply(z_pert)
    np.testing.assert_allclose(div, resid, rtol=0, atol=1e-11 * np.abs(src).max())
    assert np.abs(resid).max() > 1e-6 * np.abs(src).max()
    assert orig_tol > 0


# ====================================================================== K8

def test_k8_uniform_flow_is_a_fixed_point(oracle):
    """fluid_2d.cc:253-261: uniform face and cell velocities carry no
    advective change, so c = u + dt * force exactly."""
    p = mkparams(12, 10, bx=1.0, by=0.9)
    G, s = Grid(p), sim(oracle, p)
    U, W = 0.7, -0.4
    one = np.ones_like(G.XC)
    for aid, val in ((oracle.U, U), (oracle.V, W), (oracle.UL, U), (oracle.UD, U), (oracle.VL, W), (oracle.VD, W)):
        s.set(aid, val * one)
    fx, fy = 0.3 * np.sin(G.XC), -0.2 * G.YC
    s.set(oracle.C0, fx)
    s.set(oracle.C1, fy)
    dt = 0.05
    s.run_pass(oracle.PASS_K8, dt)
    np.testing.assert_allclose(G.inner(pad(s, oracle.C0)), G.inner(U + dt * fx), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(G.inner(pad(s, oracle.C1)), G.inner(W + dt * fy), rtol=1e-14, atol=1e-15)


# ====================================================================== boundary conditions

def test_viscous_bc_linear_shear_flow_is_exact(oracle):
    """fluid_2d_sio.cc:307-320: a linear shear u = g y driven by the stress
    tau = mu g at the interface continues linearly into both ghost rows."""
    p = mkparams(16, 12, bx=1.0, by=0.9, mu=0.37, x_sym=0)
    G, s = Grid(p), sim(oracle, p)
    g = 2.3
    s.set(oracle.U, g * G.YC)
    s.set(oracle.GTAU, np.full(p.m + 4, p.mu * g))
    s.run_pass(oracle.PASS_VISCOUS_BC)
    u = pad(s, oracle.U)
    np.testing.assert_allclose(u[1], g * G.YC[1], rtol=1e-13)
    np.testing.assert_allclose(u[0], g * G.YC[0], rtol=1e-13)


@pytest.mark.parametrize("x_sym", [0, 1])
def test_viscous_bc_ghosts_satisfy_stress_and_incompressibility(oracle, x_sym):
    """fluid_2d_sio.cc:294-343 (in its order): the mirror ghosts at x = 0
    (u odd, v even) with x_sym; then the bottom ghosts of u, whose centred
    differences across y = 0 give tau/mu of their own column (tau entry
    i + 2 for column i); then the bottom ghosts of v, which make the discrete
    divergence of each ghost row zero."""
    p = mkparams(16, 12, bx=1.0, by=0.9, mu=0.37, x_sym=x_sym)
    G, s = Grid(p), sim(oracle, p)
    rng = np.random.default_rng(9)
    u0, v0 = rng.standard_normal(G.XC.shape), rng.standard_normal(G.XC.shape)
    s.set(oracle.U, u0)
    s.set(oracle.V, v0)
    tau = rng.standard_normal(p.m + 4)
    s.set(oracle.GTAU, tau)
    s.run_pass(oracle.PASS_VISCOUS_BC)
    u, v = pad(s, oracle.U), pad(s, oracle.V)
    m, n, dx, dy = p.m, p.n, G.dx, G.dy
    rows = slice(2, n + 2)
    if x_sym:
        np.testing.assert_array_equal(u[rows, 1], -u0[rows, 2])
        np.testing.assert_array_equal(u[rows, 0], -u0[rows, 3])
        np.testing.assert_array_equal(v[rows, 1], v0[rows, 2])
        np.testing.assert_array_equal(v[rows, 0], v0[rows, 3])
    else:
        np.testing.assert_array_equal(u[rows, 0:2], u0[rows, 0:2])
        np.testing.assert_array_equal(v[rows, 0:2], v0[rows, 0:2])
    # stress: columns -2 .. m+1 (array index c+2 = column index into tau)
    np.testing.assert_allclose((u[2] - u[1]) / dy, tau / p.mu, rtol=1e-12, atol=1e-12 * np.abs(u0).max() / dy)
    np.testing.assert_allclose((u[3] - u[0]) / (3 * dy), tau / p.mu, rtol=1e-12,
                               atol=1e-12 * np.abs(u0).max() / dy)
    # incompressibility in each ghost row, columns -1 .. m-1
    cols = slice(1, m + 2)
    for row, above in ((1, 2), (0, 1)):
        dvdy = (v[above, cols] - v[row, cols]) / dy
        dudx = (u[row, 2:m + 3] - u[row, 0:m + 1]) / (2 * dx)
        np.testing.assert_allclose(dvdy + dudx, 0, atol=1e-11 * (np.abs(dvdy).max() + 1))


def test_inviscid_bc_adds_the_far_field_integral(oracle):
    """fluid_2d_sio.cc:256-290, 366-405: each right and top ghost cell gains
    dt/(pi rho) times the far-field integral of the gas pressure (closed form
    for constant pressure), accumulated onto its previous value; without
    x_sym the left ghosts mirror the right ones (u odd, v even)."""
    for x_sym in (1, 0):
        ax = 0.0 if x_sym else -1.0
        p = mkparams(128, 64, ax=ax, bx=1.0, by=0.5 if x_sym else 1.0, rho=2.0, x_sym=x_sym)
        G, s = Grid(p), sim(oracle, p)
        p0, dt = 3.0, 0.01
        s.set(oracle.GP, np.full(p.m + 1, p0))
        base = np.random.default_rng(1).standard_normal(G.XC.shape)
        s.set(oracle.U, base)
        s.set(oracle.V, 2 * base)
        s.run_pass(oracle.PASS_INVISCID_BC, dt)
        du, dv = pad(s, oracle.U) - base, pad(s, oracle.V) - 2 * base
        m, n = p.m, p.n
        pts = [(m, j) for j in range(16, n)] + [(m + 1, j) for j in range(16, n)] + \
              [(i, n) for i in range(-2, m + 2)] + [(i, n + 1) for i in range(-2, m + 2)]
        scale = 0.0
        errs = []
        for i, j in pts:
            eu, ev = (dt / (np.pi * p.rho) * q for q in
                      __import__("closed_forms").simpson_exact(G.at(G.XC, i, j), G.at(G.YC, i, j), p.ax, p.bx,
                                                               x_sym, p0))
            errs.append((G.at(du, i, j) - eu, G.at(dv, i, j) - ev))
            scale = max(scale, abs(eu), abs(ev))
        assert np.abs(np.array(errs)).max() < 1e-5 * scale
        # interior cells untouched
        np.testing.assert_array_equal(du[2:n + 2, 2:m + 2], 0)
        if not x_sym:
            u, v = pad(s, oracle.U), pad(s, oracle.V)
            np.testing.assert_array_equal(u[2:n + 2, 1], -u[2:n + 2, m + 2])
            np.testing.assert_array_equal(u[2:n + 2, 0], -u[2:n + 2, m + 3])
            np.testing.assert_array_equal(v[2:n + 2, 1], v[2:n + 2, m + 2])
            np.testing.assert_array_equal(v[2:n + 2, 0], v[2:n + 2, m + 3])


# ====================================================================== viscous solves

def _helmholtz_full(z, ghosts, vfx, vfy):
    """(I - vfx dxx - vfy dyy) z on m x n cells with explicit ghost values
    (left, right, bottom, top arrays)."""
    zl, zr, zb, zt = ghosts
    Z = np.pad(z, 1)
    Z[1:-1, 0], Z[1:-1, -1], Z[0, 1:-1], Z[-1, 1:-1] = zl, zr, zb, zt
    c = Z[1:-1, 1:-1]
    return c - vfx * (Z[1:-1, :-2] - 2 * c + Z[1:-1, 2:]) - vfy * (Z[:-2, 1:-1] - 2 * c + Z[2:, 1:-1])


@pytest.mark.parametrize("x_sym", [1, 0])
@pytest.mark.parametrize("comp", ["u", "v"])
def test_visco_rhs_and_solve_reproduce_the_implicit_step(oracle, comp, x_sym):
    """fluid_2d.cc:272-334, visco_impl.hh:28-53, visco_impl.cc:21-39: with
    c set to (I - nu dt/2 lap) z for a smooth z whose ghosts carry the
    boundary data (Dirichlet ghost values right, top and, without x_sym,
    left; the mirror at x = 0 with x_sym; the Neumann stress mean
    (tau_old + tau)/2 / mu for u and -du/dx for v at the bottom), the
    right-hand side pass and the Helmholtz solve return z."""
    m, n = 32, 24
    rho, mu, dt = 1.0, 0.3, 0.02
    ax = 0.0 if x_sym else -1.0
    p = mkparams(m, n, ax=ax, bx=1.0, by=0.8, rho=rho, mu=mu, x_sym=x_sym, implicit=1, dt=dt)
    G, s = Grid(p), sim(oracle, p)
    dx, dy = G.dx, G.dy
    vfx, vfy = mu / (2 * rho) * dt / dx ** 2, mu / (2 * rho) * dt / dy ** 2
    X, Y = G.XC, G.YC
    B = 0.45                                     # u row 0 = A + B x: -du/dx = -B
    if comp == "u":
        zf = np.sin(1.3 * X) * np.cos(0.9 * Y) + 0.2 * X * Y  # odd in x
        g = 0.37 + 0.1 * G.XC[2, 2:m + 2]                       # du/dy at y=0 per column
    else:
        zf = np.cos(1.1 * X) * (1 + 0.5 * Y * Y) + 0.3 * X * X  # even in x
        g = np.full(m, -B)
    z = G.inner(zf)
    left = (-z[:, 0] if comp == "u" else z[:, 0]) if x_sym else zf[2:n + 2, 1]
    ghosts = (left, zf[2:n + 2, m + 2], z[0] - dy * g, zf[n + 2, 2:m + 2])
    c = _helmholtz_full(z, ghosts, vfx, vfy)
    field = np.zeros_like(X)
    field[2:n + 2, 2:m + 2] = c
    s.set(oracle.C0 if comp == "u" else oracle.C1, field)
    w = zf.copy()                                 # the velocity component, ghosts included
    if comp == "v":
        uu = np.zeros_like(X)
        uu[2] = 0.2 + B * X[2]
        s.set(oracle.U, uu)
    s.set(oracle.U if comp == "u" else oracle.V, w)
    if comp == "v":
        uu = pad(s, oracle.U)
        uu[2] = 0.2 + B * X[2]
        s.set(oracle.U, uu)
    tau = np.zeros(m + 4)
    tau_old = np.zeros(m + 4)
    if comp == "u":
        tau[2:m + 2] = 1.5 * mu * g
        tau_old[2:m + 2] = 0.5 * mu * g
    s.set(oracle.GTAU, tau)
    s.set(oracle.GTAU_OLD, tau_old)
    s.run_pass(oracle.PASS_VISCO_RHS_U if comp == "u" else oracle.PASS_VISCO_RHS_V, dt)
    kind = oracle.KIND_VX if comp == "u" else oracle.KIND_VY
    s.visco_setup_and_solve(kind, vfx, vfy, 1.0)
    zs = s.get(oracle.Z_VX if comp == "u" else oracle.Z_VY).reshape(n, m)
    np.testing.assert_allclose(zs, z, rtol=0, atol=1e-11 * np.abs(z).max())


def test_visco_copy_moves_the_solutions(oracle):
    p = mkparams(12, 10, implicit=1)
    G, s = Grid(p), sim(oracle, p)
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal(p.m * p.n), rng.standard_normal(p.m * p.n)
    s.set(oracle.Z_VX, a)
    s.set(oracle.Z_VY, b)
    s.run_pass(oracle.PASS_VISCO_COPY)
    np.testing.assert_array_equal(G.inner(pad(s, oracle.C0)).ravel(), a)
    np.testing.assert_array_equal(G.inner(pad(s, oracle.C1)).ravel(), b)


# ====================================================================== K14, K15

def _node_areas(m, n, dx, dy):
    """Control area of each node of rows 1..n (row 0 is the Dirichlet row)."""
    wx = np.ones(m + 1)
    wx[0] = wx[-1] = 0.5
    wy = np.ones(n)
    wy[-1] = 0.5
    return np.outer(wy, wx) * dx * dy


def _fem_acc_bottom(oracle, p, G):
    """Diagonal of the FEM operator's Dirichlet row, from the pinned operator
    applied to unit vectors."""
    M, N = p.m + 1, p.n + 1
    A = oracle.OracleMG(oracle.KIND_FEM, M, N, G.dx, G.dy, p.x_sym)
    out = np.empty(M)
    for i in range(M):
        e = np.zeros(M * N)
        e[i] = 1
        out[i] = A.apply(e)[i]
    return out


@pytest.mark.parametrize("x_sym", [0, 1])
def test_k14_weak_divergence_uniform_and_linear(oracle, x_sym):
    """fluid_2d.cc:338-377.  The nodal right-hand side is rho/dt times the
    weak divergence: zero for a uniform field whose ghosts continue it, and
    -(rho/dt) div(c) times each node's control area (full, half on the sides
    and top, a quarter at the top corners) for a linear field whose ghosts
    continue it; the bottom row is the FEM Dirichlet row times the gas
    pressure."""
    ax = 0.0 if x_sym else -1.0
    p = mkparams(16, 12, ax=ax, bx=1.0, by=0.9, rho=997.0, x_sym=x_sym)
    G, s = Grid(p), sim(oracle, p)
    m, n, dt = p.m, p.n, 2e-4
    rdt = p.rho / dt
    gp = np.random.default_rng(4).standard_normal(m + 1) * 10
    s.set(oracle.GP, gp)
    acc = _fem_acc_bottom(oracle, p, G)
    for a, b, U, W in ((0.0, 0.0, 0.0 if x_sym else 0.6, -0.35), (1.7, -0.8, 0.0, 0.0)):
        c0 = U + a * G.XC
        c1 = W + b * G.YC
        s.set(oracle.C0, c0)
        s.set(oracle.C1, c1)
        s.set(oracle.U, c0)         # ghosts continue the field
        s.set(oracle.V, c1)
        s.run_pass(oracle.PASS_K14, dt)
        src = s.get(oracle.SRC).reshape(n + 1, m + 1)
        expect = -rdt * (a + b) * _node_areas(m, n, G.dx, G.dy)
        scale = rdt * (abs(U) + abs(W) + abs(a) + abs(b)) * G.dx
        np.testing.assert_allclose(src[1:], expect, rtol=0, atol=1e-12 * scale)
        np.testing.assert_allclose(src[0], acc * gp, rtol=1e-14)


def test_k15_linear_pressure_gives_the_exact_gradient(oracle):
    """fluid_2d.cc:383-396: u = c0 - dt/rho dp/dx at the cell centre, exact
    for bilinear nodal p; the nodal p is copied into the pressure array with
    its last column and row."""
    p = mkparams(16, 12, bx=1.0, by=0.9, rho=997.0)
    G, s = Grid(p), sim(oracle, p)
    m, n, M = p.m, p.n, p.m + 1
    al, be, ga = 3.0, -2.0, 1.5
    xn, yn = G.XN[2:n + 3, 2:m + 3], G.YN[2:n + 3, 2:m + 3]
    z = al * xn + be * yn + ga * xn * yn
    s.set(oracle.Z_FEM, z.ravel())
    rng = np.random.default_rng(5)
    c0, c1 = rng.standard_normal(G.XC.shape), rng.standard_normal(G.XC.shape)
    s.set(oracle.C0, c0)
    s.set(oracle.C1, c1)
    dt = 1e-3
    s.run_pass(oracle.PASS_K15, dt)
    X, Y = G.inner(G.XC), G.inner(G.YC)
    np.testing.assert_allclose(G.inner(pad(s, oracle.U)), G.inner(c0) - dt / p.rho * (al + ga * Y), rtol=1e-13,
                               atol=1e-15)
    np.testing.assert_allclose(G.inner(pad(s, oracle.V)), G.inner(c1) - dt / p.rho * (be + ga * X), rtol=1e-13,
                               atol=1e-15)
    np.testing.assert_array_equal(pad(s, oracle.P)[2:n + 3, 2:M + 2], z)


# ====================================================================== gas layer pieces

def _gas_arrays(oracle, p, h, pg):
    s = sim(oracle, p)
    s.set(oracle.H, h)
    s.set(oracle.PG, pg)
    return s


@pytest.mark.parametrize("x_sym", [1, 0])
def test_liquid_pressure_of_the_parabolic_interface_is_the_gas_pressure(oracle, x_sym):
    """gas_layer_solve.cc:229-249: p = pg + sigma (curvature - 1/R).  The
    initial interface h = h0 + x^2/(2R) has curvature 1/R, so p = pg on every
    node (with x_sym through the mirrored stencil at x = 0)."""
    cfg = PhysicalConfig(m=64, n=32, x_sym=bool(x_sym))
    p = cfg.params()
    dx = (p.bx - p.ax) / p.m
    x = p.ax + (np.arange(p.m) + 0.5) * dx
    h = p.h0 + x * x / (2 * p.R)
    pg = np.random.default_rng(6).standard_normal(p.m + 1) * 100
    s = _gas_arrays(oracle, p, h, pg)
    s.run_pass(oracle.PASS_GAS_LIQUID_FIELDS)
    gp = s.get(oracle.GP)
    sc = p.sigma / p.R
    if x_sym:
        np.testing.assert_allclose(gp, pg, rtol=0, atol=1e-9 * sc)
    else:
        # without x_sym the two outer nodes at each end carry pg alone
        np.testing.assert_allclose(gp[2:-2], pg[2:-2], rtol=0, atol=1e-9 * sc)
        np.testing.assert_array_equal(gp[[0, 1, -2, -1]], pg[[0, 1, -2, -1]])


def test_liquid_pressure_curvature_converges_and_mirror_stencil(oracle):
    """The curvature term is sigma h'' at the nodes to O(dx^2) for a general
    smooth interface, and with x_sym its first two nodes are the interior
    stencil applied to the interface mirrored at x = 0."""
    R, sigma = 1e-3, 0.02
    errs = []
    for m in (64, 128, 256):
        p = mkparams(m, 8, ax=0.0, bx=1e-3, by=1e-4, x_sym=1, R=R, sigma=sigma, h0=5e-6)
        dx = 1e-3 / m
        x = (np.arange(m) + 0.5) * dx
        k = 2 * np.pi / 1e-3

        def hf(x):
            return 5e-6 + 1e-6 * np.cos(k * x)
        s = _gas_arrays(oracle, p, hf(x), np.zeros(m + 1))
        s.run_pass(oracle.PASS_GAS_LIQUID_FIELDS)
        gp = s.get(oracle.GP)
        xn = np.arange(m + 1) * dx
        exact = sigma * (-1e-6 * k * k * np.cos(k * xn)) - sigma / R
        errs.append(np.abs(gp[:m - 1] - exact[:m - 1]).max())
        hm = np.concatenate([hf(x)[1::-1], hf(x)])   # mirrored: h[-2], h[-1], h[0], ...
        fac = 0.5 * sigma / dx ** 2
        for i in (0, 1):
            q = i + 2
            stencil = fac * (hm[q + 1] - hm[q] - hm[q - 1] + hm[q - 2]) - sigma / R
            np.testing.assert_allclose(gp[i], stencil, rtol=1e-12)
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 1.8), (errs, rates)


def test_gas_shear_is_the_lubrication_stress(oracle):
    """tau_i = -mu_g h_i dp/dx / 2 at cell i (the reference's formula,
    gas_layer_solve.cc:248), written at index i (quirk Q1): exact for a
    quadratic pressure, whose node difference is the derivative at the
    cell centre."""
    p = mkparams(32, 8, ax=0.0, bx=2.0, by=0.5, x_sym=1, sigma=0.0, mug=1.8e-5)
    m = p.m
    dx = 2.0 / m
    xn = np.arange(m + 1) * dx
    xc = (np.arange(m) + 0.5) * dx
    h = 1.0 + 0.3 * xc
    s = _gas_arrays(oracle, p, h, 4.0 * xn * xn)
    s.run_pass(oracle.PASS_GAS_LIQUID_FIELDS)
    tau = s.get(oracle.GTAU)
    np.testing.assert_allclose(tau[:m], -0.5 * p.mug * h * (8.0 * xc), rtol=1e-12)


def test_tau_boundary_condition_reflects_oddly(oracle):
    """gas_layer.cc:25-30: tau = 0 at both ends by odd reflection of the two
    outer entries."""
    p = mkparams(16, 8)
    s = sim(oracle, p)
    t = np.random.default_rng(8).standard_normal(p.m + 4)
    s.set(oracle.GTAU, t)
    s.run_pass(oracle.PASS_GAS_TAU_BC)
    r = s.get(oracle.GTAU)
    ml = p.m + 4
    assert r[0] == -t[3] and r[1] == -t[2] and r[ml - 2] == -t[ml - 3] and r[ml - 1] == -t[ml - 4]
    np.testing.assert_array_equal(r[2:ml - 2], t[2:ml - 2])


@pytest.mark.parametrize("x_sym", [1, 0])
def test_height_slope_exact_for_quadratic_interface(oracle, x_sym):
    """gas_layer_solve.cc:58-62: hx at the interior nodes is the centred
    difference (exact for a quadratic); zero at the symmetry line with x_sym,
    copied from the neighbour at the outer ends otherwise."""
    p = mkparams(32, 8, ax=0.0 if x_sym else -1.0, bx=1.0, x_sym=x_sym)
    m = p.m
    dx = (p.bx - p.ax) / m
    xc = p.ax + (np.arange(m) + 0.5) * dx
    xn = p.ax + np.arange(m + 1) * dx
    s = _gas_arrays(oracle, p, 2.0 + 0.7 * xc + 1.3 * xc * xc, np.zeros(m + 1))
    s.run_pass(oracle.PASS_GAS_HX)
    hx = s.get(oracle.HX)
    np.testing.assert_allclose(hx[1:m], 0.7 + 2.6 * xn[1:m], rtol=1e-12)
    assert hx[0] == (0.0 if x_sym else hx[1])
    assert hx[m] == hx[m - 1]


def test_gas_initial_state_is_a_falling_parabola_at_ambient_pressure(oracle):
    """gas_layer_solve.cc:32-53: h = h0 + x^2/(2R) at the cell centres, the
    interface falls at V, the gas is at ambient pressure (pg = 0), and the
    liquid pressure and shear start at zero (curvature 1/R balances
    sigma/R)."""
    p = PhysicalConfig(m=64, n=32).params()
    s = sim(oracle, p)
    s.run_pass(oracle.PASS_GAS_INIT)
    dx = (p.bx - p.ax) / p.m
    x = p.ax + (np.arange(p.m) + 0.5) * dx
    np.testing.assert_allclose(s.get(oracle.H), p.h0 + x * x / (2 * p.R), rtol=1e-15)
    assert np.all(s.get(oracle.HT) == -p.V)
    assert not np.any(s.get(oracle.PG)) and not np.any(s.get(oracle.FBD))
    sc = p.sigma / p.R
    assert np.abs(s.get(oracle.GP)).max() < 1e-9 * sc
    assert np.abs(s.get(oracle.GTAU)[:p.m]).max() < 1e-9 * sc * p.mug * p.h0 / dx


# ====================================================================== initial guess

@pytest.mark.parametrize("which", ["mac", "fem"])
def test_initial_guess_is_exact_for_solutions_linear_in_time(oracle, which):
    """mgs_common.cc:17-50: after two solves, the guess at time t is the
    linear extrapolation through the last two solutions, so it is exact when
    the solution is linear in time (and the history moves on)."""
    p = PhysicalConfig(m=32, n=16).params()
    s = sim(oracle, p)
    kind = oracle.KIND_MAC if which == "mac" else oracle.KIND_FEM
    zid = oracle.Z_MAC if which == "mac" else oracle.Z_FEM
    N = s.get(zid).size
    rng = np.random.default_rng(12)
    A, B = rng.standard_normal(N), rng.standard_normal(N)

    def zt(t):
        return A + B * t
    times = [0.1, 0.35, 0.6, 1.1, 1.2]
    for k, t in enumerate(times):
        s.initial_guess(kind, t)
        if k >= 2:
            np.testing.assert_allclose(s.get(zid), zt(t), rtol=1e-12, atol=1e-12)
        s.set(zid, zt(t))        # the solve's result at t


def test_initial_guess_skips_duplicate_times(oracle, capfd):
    """Two solves at the same time leave the history unable to extrapolate:
    the guess keeps the last solution and a warning is printed
    (mgs_common.cc:32-35)."""
    p = PhysicalConfig(m=32, n=16).params()
    s = sim(oracle, p)
    z = np.arange(32 * 16, dtype=float)
    for t in (0.5, 0.5):
        s.initial_guess(oracle.KIND_MAC, t)
        s.set(oracle.Z_MAC, z)
    capfd.readouterr()
    s.initial_guess(oracle.KIND_MAC, 0.7)
    np.testing.assert_array_equal(s.get(oracle.Z_MAC), z)
    assert "duplicate times" in capfd.readouterr().err


# ====================================================================== mirror symmetry

def _half_and_full(oracle, m, n, steps, q1, q13):
    half = PhysicalConfig(m=m, n=n, x_sym=True).params()
    full = PhysicalConfig(m=2 * m, n=n, x_sym=False).params()
    assert abs((full.bx - full.ax) / full.m - (half.bx - half.ax) / half.m) < 1e-18
    assert abs(full.by - half.by) < 1e-18 and full.dt_reg == half.dt_reg
    a = oracle.OracleSim(half, rel_tol=1e-13)
    b = oracle.OracleSim(full, rel_tol=1e-13)
    for s in (a, b):
        s.set_quirks(q1, q13)
        s.init_fields()
    for _ in range(steps):
        a.step_forward(half.dt_reg)
        b.step_forward(full.dt_reg)
    R = slice(2, n + 2)

    def e(x, y):
        return np.linalg.norm(x - y) / np.linalg.norm(y)
    ua, va, pa = a.padded(oracle.U), a.padded(oracle.V), a.padded(oracle.P)
    ub, vb, pb = b.padded(oracle.U), b.padded(oracle.V), b.padded(oracle.P)
    right = lambda x: x[R, m + 2:2 * m + 2]             # right half of the full run's cells
    left = lambda x: x[R, 2:m + 2][:, ::-1]             # left half, mirrored
    return {"u": e(right(ub), ua[R, 2:m + 2]), "v": e(right(vb), va[R, 2:m + 2]),
            "p": e(pb[2:n + 3, m + 2:2 * m + 3], pa[2:n + 3, 2:m + 3]),
            "h": e(b.get(oracle.H)[m:], a.get(oracle.H)), "pg": e(b.get(oracle.PG)[m:], a.get(oracle.PG)),
            "mirror u": e(-left(ub), right(ub)), "mirror v": e(left(vb), right(vb))}


def test_symmetric_half_domain_equals_full_width_run(oracle):
    """The whole step with x_sym on [0, L] (mirror conditions at x = 0: odd
    u, even v ghosts, the image terms of the far-field integral, the
    symmetric Newton row of the gas layer, zero normal face velocity, no
    flux term at x = 0) against the full-width step on [-L, L] (none of that
    code) from the same mirror-symmetric state, six steps, solves to 1e-13.
    Two reference quirks break the mirror symmetry of the full-width step:
    the interface shear read two cells off (Q1, a shift in +x) and the height
    update sweeping in place from left to right (Q13).  With both switched
    off in the oracle, the half run equals the right half of the full run to
    rounding and the full run is mirror-symmetric to rounding; with them on
    (the reference's behaviour) the difference stays at the size those
    quirks alone produce."""
    d = _half_and_full(oracle, 32, 16, 6, 0, 0)
    for k, err in d.items():
        assert err < 1e-12, (k, d)
    d = _half_and_full(oracle, 32, 16, 6, 1, 1)
    for k, err in d.items():
        assert err < 1e-4, (k, d)


# ====================================================================== tracers

def test_tracers_move_with_the_bilinear_velocity(oracle):
    """fluid_2d.cc:507-525: each tracer moves by dt times the velocity
    interpolated bilinearly between the four cell centres around it.  Exact
    for a field linear in x and y everywhere (also for tracers outside the
    cell centres, which extrapolate from the edge cells); for a field with a
    kink along a column of cell centres, exact only when the cell to the
    lower left is the one used."""
    p = mkparams(16, 12, bx=1.0, by=0.9)
    G = Grid(p)
    rng = np.random.default_rng(13)
    pos = np.column_stack([rng.uniform(-0.05, 1.05, 200), rng.uniform(-0.05, 0.95, 200)])
    dt = 0.01
    s = sim(oracle, p)
    s.set(oracle.U, 0.3 + 0.7 * G.XC - 0.4 * G.YC)
    s.set(oracle.V, -0.2 + 0.5 * G.XC + 0.25 * G.YC)
    s.set_tracers(pos)
    s.run_pass_tracers(dt)
    t = s.tracers()
    x, y = pos[:, 0], pos[:, 1]
    np.testing.assert_allclose(t[:, 0], x + dt * (0.3 + 0.7 * x - 0.4 * y), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(t[:, 1], y + dt * (-0.2 + 0.5 * x + 0.25 * y), rtol=1e-14, atol=1e-15)
    # kink at the centre of column 7: |x - x_7| is linear between neighbouring
    # centres, so tracers strictly inside the domain move exactly
    xk = G.at(G.XC, 7, 0)
    s2 = sim(oracle, p)
    s2.set(oracle.U, np.abs(G.XC - xk))
    inside = np.column_stack([rng.uniform(G.dx, 1 - G.dx, 300), rng.uniform(G.dy, 0.9 - G.dy, 300)])
    s2.set_tracers(inside)
    s2.run_pass_tracers(dt)
    np.testing.assert_allclose(s2.tracers()[:, 0], inside[:, 0] + dt * np.abs(inside[:, 0] - xk), rtol=1e-14,
                               atol=1e-15)
