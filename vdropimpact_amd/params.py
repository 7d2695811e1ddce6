"""Simulation parameters for the per-timestep fluid solve.

``VdiParams`` is the ctypes image of ``vdi_params`` (include/vdi.h).  The
derivation from physical inputs follows the reference's configuration
reader: fileinfo.cc:141-152 (viscosity, the non-dimensional scalings with
st = mug/(rho V R), the domain) and fluid_2d.cc:15-17 (grid spacing and the
regular timestep).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass


class VdiParams(ctypes.Structure):
    """Field-for-field image of ``vdi_params`` in include/vdi.h."""

    _fields_ = [
        ("m", ctypes.c_int), ("n", ctypes.c_int),
        ("x_sym", ctypes.c_int), ("implicit_visc", ctypes.c_int),
        ("ax", ctypes.c_double), ("bx", ctypes.c_double),
        ("ay", ctypes.c_double), ("by", ctypes.c_double),
        ("rho", ctypes.c_double), ("mu", ctypes.c_double),
        ("dt_reg", ctypes.c_double),
        ("alpha", ctypes.c_double), ("mug", ctypes.c_double), ("Pamb", ctypes.c_double),
        ("h0", ctypes.c_double), ("R", ctypes.c_double), ("V", ctypes.c_double),
        ("sigma", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


@dataclass
class PhysicalConfig:
    """The keywords of a reference ``.cfg`` file that the step depends on
    (fileinfo.cc:52-120).  Units are SI, as in the reference."""

    m: int
    n: int
    L_nd: float = 30.0
    h0_nd: float = 4.586010
    nul: float = 32e-6          # kinematic viscosity, m^2/s (nul_cSt * 1e-6)
    rhol: float = 1000.0
    R: float = 1e-3
    V: float = 0.5
    gamma: float = 1.4          # alpha = 1/gamma
    sigma: float = 0.02
    mug: float = 1.8e-5
    Pamb: float = 1e5
    tmult: float = 8e-3
    x_sym: bool = True
    implicit_visc: bool = True

    def params(self) -> VdiParams:
        """Derived constants, fileinfo.cc:141-152 and fluid_2d.cc:15-17."""
        if self.m <= 0 or self.n <= 0:
            raise ValueError("Grid dimensions must be positive")
        rho = self.rhol
        mu = self.nul * rho
        st = self.mug / (rho * self.V * self.R)
        st3 = st ** (1.0 / 3.0)
        st23 = st3 * st3
        L = self.L_nd * self.R * st3
        h0 = self.h0_nd * self.R * st23
        bx, by = L, L * self.n / self.m
        if self.x_sym:
            ax = 0.0
        else:
            ax = -L
            by *= 2
        dx = (bx - ax) / self.m
        if self.implicit_visc:
            dt_reg = dx * self.tmult
        else:
            dt_reg = dx * dx * rho * (1.0 / mu) * self.tmult * 0.25
        return VdiParams(
            m=self.m, n=self.n, x_sym=int(self.x_sym), implicit_visc=int(self.implicit_visc),
            ax=ax, bx=bx, ay=0.0, by=by, rho=rho, mu=mu, dt_reg=dt_reg,
            alpha=1.0 / self.gamma, mug=self.mug, Pamb=self.Pamb, h0=h0, R=self.R,
            V=self.V, sigma=self.sigma,
        )


def drop_impact_params(m: int, n: int, **overrides) -> VdiParams:
    """The single-drop workload: nu=32 cSt, rho=1000 kg/m^3, sigma=0.02 N/m,
    V=0.5 m/s, R=1 mm, h_min=5 um (h0_nd=4.586010 with L_nd=30), gas viscosity
    1.8e-5 Pa s, x_sym, implicit viscosity, tmult=8e-3."""
    return PhysicalConfig(m=m, n=n, **overrides).params()


def h0_nd_for_hmin(h_min: float, mug: float, rho: float, V: float, R: float) -> float:
    """Inverse of h0 = h0_nd R st^(2/3) (fileinfo.cc:145-146)."""
    st = mug / (rho * V * R)
    return h_min / (R * st ** (2.0 / 3.0))


__all__ = ["VdiParams", "PhysicalConfig", "drop_impact_params", "h0_nd_for_hmin"]
