"""Closed-form eigenpairs and integrals that pin the operators independently
of any implementation.

Each operator is separable, so tensor products of 1D eigenvectors chosen to
satisfy the boundary conditions are exact eigenvectors:

MAC (mgs_mac.hh:37-50), m x n cells, A = xxsp d_xx + yysp d_yy:
  x Neumann both sides:  cos(k pi (i+1/2)/m),        lam = -4 xxsp sin^2(k pi/2m)
  y Dirichlet face j=-1/2, Neumann top:
                         sin((l+1/2) pi (j+1/2)/n),  lam = -4 yysp sin^2((l+1/2) pi/2n)

visco (visco_impl.hh:32-53), A = I - vfx d_xx - vfy d_yy:
  y Neumann bottom, Dirichlet ghost centre at j=n:
                         cos((l+1/2) pi (j+1/2)/(n+1/2))
  x, x_sym, u: odd face at i=-1/2, ghost centre i=m zero: sin(k pi (i+1/2)/(m+1/2))
  x, x_sym, v: even face at i=-1/2:                      cos((k+1/2) pi (i+1/2)/(m+1/2))
  x, no x_sym: ghost centres i=-1 and i=m zero:          sin(k pi (i+1)/(m+1))
  lam = 1 + 4 vfx sin^2(tx/2) + 4 vfy sin^2(ty/2)

FEM (mgs_fem.hh:42-56, mgs_fem.cc:34-50), M x N nodes: the bilinear stiffness
  K = (dy/dx) Kx (x) My + (dx/dy) Mx (x) Ky with K1 = tridiag(-1,2,-1),
  M1 = tridiag(1,4,1)/6, natural-boundary end rows halved.  For
  phi = cos(k pi i/(M-1)) sin((l+1/2) pi j/(N-1)) (zero on the Dirichlet row):
  (A phi)_ij = lam * wx_i * wy_j * phi_ij, wx = 1/2 at i in {0,M-1}, wy = 1/2
  at j = N-1, and lam = (dy/dx)(2-2cx)(4+2cy)/6 + (dx/dy)(4+2cx)(2-2cy)/6.
"""
from __future__ import annotations

import numpy as np

MAC, FEM, VX, VY = 0, 1, 2, 3


def mac_mode(m, n, k, l, xxsp, yysp):
    i = np.arange(m)
    j = np.arange(n)
    phx = np.cos(k * np.pi * (i + 0.5) / m)
    phy = np.sin((l + 0.5) * np.pi * (j + 0.5) / n)
    lam = -4 * xxsp * np.sin(k * np.pi / (2 * m)) ** 2 - 4 * yysp * np.sin((l + 0.5) * np.pi / (2 * n)) ** 2
    z = np.outer(phy, phx)
    return z.ravel(), (lam * z).ravel()


def visco_mode(m, n, k, l, vfx, vfy, x_sym, horizontal):
    i = np.arange(m)
    j = np.arange(n)
    ty = (l + 0.5) * np.pi / (n + 0.5)
    phy = np.cos(ty * (j + 0.5))
    if x_sym and horizontal:
        tx = k * np.pi / (m + 0.5)
        phx = np.sin(tx * (i + 0.5))
    elif x_sym:
        tx = (k + 0.5) * np.pi / (m + 0.5)
        phx = np.cos(tx * (i + 0.5))
    else:
        tx = k * np.pi / (m + 1)
        phx = np.sin(tx * (i + 1))
    lam = 1 + 4 * vfx * np.sin(tx / 2) ** 2 + 4 * vfy * np.sin(ty / 2) ** 2
    z = np.outer(phy, phx)
    return z.ravel(), (lam * z).ravel()


def fem_mode(M, N, k, l, dx, dy):
    i = np.arange(M)
    j = np.arange(N)
    tx = k * np.pi / (M - 1)
    ty = (l + 0.5) * np.pi / (N - 1)
    phx = np.cos(tx * i)
    phy = np.sin(ty * j)
    cx, cy = np.cos(tx), np.cos(ty)
    lam = (dy / dx) * (2 - 2 * cx) * (4 + 2 * cy) / 6 + (dx / dy) * (4 + 2 * cx) * (2 - 2 * cy) / 6
    wx = np.ones(M)
    wx[0] = wx[-1] = 0.5
    wy = np.ones(N)
    wy[-1] = 0.5
    z = np.outer(phy, phx)
    az = lam * np.outer(wy, wx) * z
    az[0, :] = 0.0
    return z.ravel(), az.ravel()


def burgers_godunov(u0, u1):
    """State at x/t=0 of the Riemann problem for u_t + (u^2/2)_x = 0 with
    left state u0, right state u1; returns (value, side) with side 0 for the
    left state, 1 for the right state and 2 for the sonic (zero) state."""
    if u0 > u1:  # shock, speed (u0+u1)/2; ties go to the right state
        return (u0, 0) if u0 + u1 > 0 else (u1, 1)
    if u0 > 0:
        return u0, 0
    if u1 < 0:
        return u1, 1
    return 0.0, 2


def simpson_exact(x, y, ax, bx, x_sym, p0):
    """Exact values of the two boundary integrals for constant pressure p0 on
    [ax,bx]:  du = 2 y p0 I1,  dv = -p0 I2 with
    I1 = int (x-s)/((x-s)^2+y^2)^2 ds,  I2 = int ((x-s)^2-y^2)/((x-s)^2+y^2)^2 ds
    (plus the image x+s terms when x_sym)."""
    def F1(s):  # antiderivative in s of (x-s)/r^4
        return 1.0 / (2 * ((x - s) ** 2 + y * y))

    def F2(s):  # antiderivative in s of ((x-s)^2-y^2)/r^4
        t = x - s
        return t / (t * t + y * y)

    I1 = F1(bx) - F1(ax)
    I2 = F2(bx) - F2(ax)
    if x_sym:
        def G1(s):
            return -1.0 / (2 * ((x + s) ** 2 + y * y))

        def G2(s):
            t = x + s
            return -t / (t * t + y * y)

        I1 += G1(bx) - G1(ax)
        I2 += G2(bx) - G2(ax)
    return 2 * y * p0 * I1, -p0 * I2
