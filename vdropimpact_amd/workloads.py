"""Synthetic inputs for the operator-level solves (projection and Helmholtz
systems alone, without the rest of the time step).

Velocity field: the sum of 64 Fourier modes with integer wavenumbers
kx, ky in 1..16, amplitudes N(0,1)/|k|^2 and random phases,
  u(x,y) = sum_q a_q / |k_q|^2 cos(2 pi (kx_q x + ky_q y) + phi_q),
with v drawn independently.  The random numbers come from SplitMix64 and a
Box-Muller transform, so every consumer sees the same bits for a seed.
Grids: m x n cells, dx = dy = 1/m.

Right-hand sides, following the reference's own formulas:
  MAC  : face divergence xsp (ul[i+1]-ul[i]) + ysp (vd[j+1]-vd[j]) with u
         sampled on x-faces and v on y-faces (fluid_2d.cc:218-226);
  FEM  : nodal weak divergence of cell-centred (u,v) with rho/dt = 1
         (fluid_2d.cc:338-360), Dirichlet row b = 0;
  Helmholtz: the cell-centred u, with vfx = vfy = 0.132.
"""
from __future__ import annotations

import numpy as np

MASK64 = (1 << 64) - 1
SEED = 20261203


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """(0, 1], 53 random bits."""
        return ((self.next_u64() >> 11) + 1) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        u1, u2 = self.uniform(), self.uniform()
        return float(np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2))

    def integer(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1) - 1e-12)


def fourier_modes(seed: int = SEED, nmodes: int = 64, kmax: int = 16):
    rng = SplitMix64(seed)
    modes = []
    for _comp in range(2):
        comp = []
        for _q in range(nmodes):
            kx, ky = rng.integer(1, kmax), rng.integer(1, kmax)
            a = rng.normal()
            phi = 2.0 * np.pi * rng.uniform()
            comp.append((kx, ky, a / (kx * kx + ky * ky), phi))
        modes.append(comp)
    return modes


def _field(comp, x, y):
    f = np.zeros(np.broadcast(x, y).shape)
    for kx, ky, amp, phi in comp:
        f += amp * np.cos(2.0 * np.pi * (kx * x + ky * y) + phi)
    return f


def velocity(m: int, n: int, seed: int = SEED):
    """Cell-centred u, v (n x m) and the face samples ul (n x (m+1)),
    vd ((n+1) x m)."""
    cu, cv = fourier_modes(seed)
    dx = dy = 1.0 / m
    xc = (np.arange(m) + 0.5) * dx
    yc = (np.arange(n) + 0.5) * dy
    xf = np.arange(m + 1) * dx
    yf = np.arange(n + 1) * dy
    u = _field(cu, xc[None, :], yc[:, None])
    v = _field(cv, xc[None, :], yc[:, None])
    ul = _field(cu, xf[None, :], yc[:, None])
    vd = _field(cv, xc[None, :], yf[:, None])
    return u, v, ul, vd


def advection_state(m: int, n: int, seed: int = SEED):
    """Padded (n+4) x (m+4) arrays for the advection passes alone: u, v at the
    cell centres (ghost cells included) from the 64-mode field, the pressure
    at the nodes (kept in the padded layout, node (i, j) at (i dx, j dy))
    from the 64 modes of seed + 1, and a smooth MAC solution z (m x n) from
    those of seed + 2; domain [0, 1] x [0, n/m]."""
    cu, cv = fourier_modes(seed)
    cp, _ = fourier_modes(seed + 1)
    cz, _ = fourier_modes(seed + 2)
    dx = 1.0 / m
    i = np.arange(-2, m + 2)
    j = np.arange(-2, n + 2)
    xc, yc = (i + 0.5) * dx, (j + 0.5) * dx
    xn, yn = i * dx, j * dx
    u = _field(cu, xc[None, :], yc[:, None])
    v = _field(cv, xc[None, :], yc[:, None])
    p = _field(cp, xn[None, :], yn[:, None])
    z = 1e-3 * _field(cz, xc[None, 2:m + 2], yc[2:n + 2, None])
    return u, v, p, z


def mac_rhs(m: int, n: int, seed: int = SEED) -> np.ndarray:
    _u, _v, ul, vd = velocity(m, n, seed)
    xsp = ysp = float(m)
    return (xsp * (ul[:, 1:] - ul[:, :-1]) + ysp * (vd[1:, :] - vd[:-1, :])).ravel()


def fem_rhs(m: int, n: int, seed: int = SEED) -> np.ndarray:
    """(n+1) x (m+1) nodes; interior weak divergence as fluid_2d.cc:338-360
    with rho/dt = 1, natural boundaries without flux terms, Dirichlet row 0."""
    u, v, _ul, _vd = velocity(m, n, seed)
    dx = dy = 1.0 / m
    me, ne = m + 1, n + 1
    spx = np.zeros((ne, me))
    spy = np.zeros((ne, me))
    # cell (i,j) contributes to its four corner nodes
    spx[1:, 1:] += u          # bottom-left cell of node (i+1, j+1)
    spy[1:, 1:] += v
    spx[:-1, 1:] += u         # top-left
    spy[:-1, 1:] -= v
    spx[1:, :-1] -= u         # bottom-right
    spy[1:, :-1] += v
    spx[:-1, :-1] -= u        # top-right
    spy[:-1, :-1] -= v
    b = 0.5 * (dy * spx + dx * spy)
    b[0, :] = 0.0
    return b.ravel()


def helmholtz_rhs(m: int, n: int, seed: int = SEED) -> np.ndarray:
    u, _v, _ul, _vd = velocity(m, n, seed)
    return u.ravel()


VISCO_COUPLING = 0.132


def operator_case(kind: int, m: int, n: int, seed: int = SEED):
    """(m_unknowns, n_unknowns, cx, cy, b) for kind 0 MAC, 1 FEM, 2/3 visco."""
    if kind == 0:
        return m, n, float(m * m), float(m * m), mac_rhs(m, n, seed)
    if kind == 1:
        return m + 1, n + 1, 1.0 / m, 1.0 / m, fem_rhs(m, n, seed)
    return m, n, VISCO_COUPLING, VISCO_COUPLING, helmholtz_rhs(m, n, seed)


BATCH_GRIDS = [(256, 128), (512, 256), (1024, 512)]


def batch_members(count: int = 64, seed: int = SEED, grids=None):
    """Parameters of the batched workload: member s draws, from SplitMix64
    seeded with seed + s, a grid uniformly from `grids` (default 256x128,
    512x256, 1024x512), a kinematic viscosity log-uniform over 1-100 cSt, an
    impact speed uniform over 0.3-1.0 m/s and a drop radius uniform over
    0.5-2 mm; h0_nd is recomputed so that h_min = 5 um.  Other physics as the
    single-drop workload.  Returns (PhysicalConfig list, VdiParams list)."""
    from .params import PhysicalConfig, h0_nd_for_hmin
    grids = grids or BATCH_GRIDS
    cfgs, prms = [], []
    for s in range(count):
        rng = SplitMix64(seed + s)
        m, n = grids[rng.integer(0, len(grids) - 1)]
        nu_cst = 10.0 ** (2.0 * rng.uniform())
        V = 0.3 + 0.7 * rng.uniform()
        R = 0.5e-3 + 1.5e-3 * rng.uniform()
        h0_nd = h0_nd_for_hmin(5e-6, 1.8e-5, 1000.0, V, R)
        cfg = PhysicalConfig(m=m, n=n, nul=nu_cst * 1e-6, V=V, R=R, h0_nd=h0_nd)
        cfgs.append(cfg)
        prms.append(cfg.params())
    return cfgs, prms


__all__ = ["SplitMix64", "fourier_modes", "advection_state", "batch_members", "BATCH_GRIDS", "velocity", "mac_rhs", "fem_rhs", "helmholtz_rhs", "operator_case",
           "SEED", "VISCO_COUPLING"]
