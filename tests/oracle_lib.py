"""ctypes binding to the CPU FP64 oracle (oracle/libvdo.so).

Test infrastructure: the oracle is the checker.  Only tests, smoke() and the
bench's cpu_baseline leg use this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.environ.get("VDO_ORACLE_SO") or os.path.join(ORACLE_DIR, "libvdo.so")

# array ids (oracle/vdo.h, same numbering as include/vdi.h)
(U, V, P, UL, VL, UD, VD, C0, C1, SRC, GP, GTAU, GP_OLD, GTAU_OLD, H, HX, HT, FBD, PG, PG_OLD,
 Z_MAC, Z_FEM, Z_VX, Z_VY) = range(24)
KIND_MAC, KIND_FEM, KIND_VX, KIND_VY = 0, 1, 2, 3
# oracle-only scratch fields of the step (oracle/vdo.h)
UX, UY, VXS, VYS, UR, VR, UU, VU, C2, C3 = range(64, 74)
# passes of the step (oracle/vdo.h, VDO_PASS_*)
(PASS_K1, PASS_K2, PASS_K3, PASS_K4, PASS_K5, PASS_K6, PASS_K7, PASS_K8, PASS_INVISCID_BC, PASS_VISCOUS_BC,
 PASS_VISCO_RHS_U, PASS_VISCO_RHS_V, PASS_VISCO_COPY, PASS_K14, PASS_K15, PASS_GAS_LIQUID_FIELDS,
 PASS_GAS_TAU_BC, PASS_GAS_HX, PASS_GAS_INIT, PASS_TRACERS) = range(1, 21)

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def build(force: bool = False) -> str:
    if os.environ.get("VDO_ORACLE_SO"):
        return ORACLE_SO          # a prebuilt variant (mutation tests)
    if force or not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return ORACLE_SO


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(ORACLE_SO)
        vp = ctypes.c_void_p
        L.vdo_create.restype = vp
        L.vdo_create.argtypes = [ctypes.c_void_p]
        L.vdo_destroy.argtypes = [vp]
        L.vdo_set_mg.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.vdo_set_schedule.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.vdo_set_schedule.restype = ctypes.c_int
        L.vdo_set_schedule_levels.argtypes = [vp] + [ctypes.c_int] * 5
        L.vdo_set_schedule_levels.restype = ctypes.c_int
        L.vdo_init_fields.argtypes = [vp]
        L.vdo_step_forward.argtypes = [vp, ctypes.c_double]
        L.vdo_gas_update.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.vdo_gas_height.argtypes = [vp, ctypes.c_double]
        L.vdo_gas_newton.argtypes = [vp, ctypes.c_double]
        L.vdo_time.restype = ctypes.c_double
        L.vdo_time.argtypes = [vp]
        L.vdo_vframe.restype = ctypes.c_double
        L.vdo_vframe.argtypes = [vp]
        L.vdo_array_len.restype = ctypes.c_long
        L.vdo_array_len.argtypes = [vp, ctypes.c_int]
        L.vdo_get_array.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_long]
        L.vdo_set_array.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_long]
        L.vdo_set_time.argtypes = [vp, ctypes.c_double]
        L.vdo_last_cycles.argtypes = [vp, ctypes.c_int]
        L.vdo_last_newton.argtypes = [vp]
        L.vdo_set_quirks.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.vdo_set_threads.argtypes = [ctypes.c_int]
        L.vdo_set_threads.restype = ctypes.c_int
        L.vdo_set_tracers.argtypes = [vp, _dp, ctypes.c_int]
        L.vdo_get_tracers.argtypes = [vp, _dp, ctypes.c_int]
        L.vdo_pass.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.vdo_pass.restype = ctypes.c_int
        L.vdo_initial_guess.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.vdo_initial_guess.restype = ctypes.c_int
        L.vdo_solve_v_cycle.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.vdo_solve_v_cycle.restype = ctypes.c_int
        L.vdo_visco_setup_and_solve.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.vdo_visco_setup_and_solve.restype = ctypes.c_int
        L.vdo_mg_create.restype = vp
        L.vdo_mg_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int]
        L.vdo_mg_destroy.argtypes = [vp]
        L.vdo_mg_set.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.vdo_mg_levels.argtypes = [vp]
        L.vdo_mg_apply.argtypes = [vp, _dp, _dp]
        L.vdo_mg_resnorm.restype = ctypes.c_double
        L.vdo_mg_resnorm.argtypes = [vp, _dp, _dp]
        L.vdo_mg_solve.argtypes = [vp, _dp, _dp, ctypes.c_double, ctypes.c_int]
        L.vdo_mg_vcycles.argtypes = [vp, _dp, _dp, ctypes.c_int]
        L.vdo_mono_diff.restype = ctypes.c_double
        L.vdo_mono_diff.argtypes = [ctypes.c_double] * 5
        L.vdo_godunov_set.argtypes = [_dp] * 4
        L.vdo_godunov_set_tang.argtypes = [_dp] * 4
        L.vdo_simpson_velocity.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           _dp, _dp]
        L.vdo_tri_solve_pivot.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.vdo_tri_solve_thomas.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class OracleSim:
    """The oracle's whole-step simulation (vdo_sim)."""

    def __init__(self, params, nu1=2, nu2=2, rel_tol=1e-10, max_cycles=200):
        self.params = params
        self.m, self.n = params.m, params.n
        self.h = lib().vdo_create(ctypes.byref(params))
        lib().vdo_set_mg(self.h, nu1, nu2, rel_tol, max_cycles)

    def set_schedule(self, which, nu1, nu2, nu1s=None, nu2s=None):
        nu1s = nu1 if nu1s is None else nu1s
        nu2s = nu2 if nu2s is None else nu2s
        if lib().vdo_set_schedule_levels(self.h, which, nu1, nu2, nu1s, nu2s) != 0:
            raise ValueError("vdo_set_schedule_levels")

    def close(self):
        if self.h:
            lib().vdo_destroy(self.h)
            self.h = None

    __del__ = close

    def init_fields(self):
        lib().vdo_init_fields(self.h)

    def step_forward(self, dt):
        rc = lib().vdo_step_forward(self.h, dt)
        if rc:
            raise RuntimeError(f"oracle step failed with status {rc}")

    def gas_update(self, t, dt):
        rc = lib().vdo_gas_update(self.h, t, dt)
        if rc:
            raise RuntimeError(f"oracle gas update failed with status {rc}")

    def gas_height(self, dt):
        lib().vdo_gas_height(self.h, dt)

    def gas_newton(self, dt):
        return lib().vdo_gas_newton(self.h, dt)

    @property
    def time(self):
        return lib().vdo_time(self.h)

    @time.setter
    def time(self, t):
        lib().vdo_set_time(self.h, t)

    def get(self, aid):
        n = lib().vdo_array_len(self.h, aid)
        out = np.empty(n, np.float64)
        assert lib().vdo_get_array(self.h, aid, ptr(out), n) == 0
        return out

    def set(self, aid, arr):
        arr = np.ascontiguousarray(arr, np.float64).ravel()
        assert lib().vdo_set_array(self.h, aid, ptr(arr), arr.size) == 0

    def padded(self, aid):
        return self.get(aid).reshape(self.n + 4, self.m + 4)

    def last_cycles(self, kind):
        return lib().vdo_last_cycles(self.h, kind)

    def set_tracers(self, pos):
        a = np.ascontiguousarray(pos, np.float64).ravel()
        self.ntrace = a.size // 2
        lib().vdo_set_tracers(self.h, ptr(a), self.ntrace)

    def tracers(self):
        nt = getattr(self, "ntrace", 0)
        out = np.empty(2 * nt, np.float64)
        assert lib().vdo_get_tracers(self.h, ptr(out), nt) == 0
        return out.reshape(nt, 2)

    def run_pass_tracers(self, dt):
        """The tracer update alone (the first thing a step does)."""
        if lib().vdo_pass(self.h, PASS_TRACERS, dt) != 0:
            raise ValueError("tracer pass")

    def set_quirks(self, q1_tau_shift=1, q13_inplace=1):
        """Test switches of the reference quirks Q1 and Q13 (1 = as in the reference)."""
        lib().vdo_set_quirks(self.h, q1_tau_shift, q13_inplace)

    def run_pass(self, pass_id, dt=0.0):
        """One pass of the step (PASS_*) on the current arrays."""
        if lib().vdo_pass(self.h, pass_id, dt) != 0:
            raise ValueError(f"pass {pass_id} does not apply")

    def initial_guess(self, which, t):
        assert lib().vdo_initial_guess(self.h, which, t) == 0

    def solve_v_cycle(self, which, t):
        """mgs_mac/mgs_fem::solve_v_cycle(t) on the shared src; V-cycles."""
        c = lib().vdo_solve_v_cycle(self.h, which, t)
        if c < 0:
            raise RuntimeError("oracle solve failed")
        return c

    def visco_setup_and_solve(self, which, vfx, vfy, t):
        c = lib().vdo_visco_setup_and_solve(self.h, which, vfx, vfy, t)
        if c < 0:
            raise RuntimeError("oracle viscous solve failed")
        return c

    def last_newton(self):
        return lib().vdo_last_newton(self.h)


class OracleMG:
    """A standalone oracle linear system (vdo_mg)."""

    def __init__(self, kind, m, n, cx, cy, x_sym=1, nu1=2, nu2=2, ncs=8):
        self.kind, self.m, self.n = kind, m, n
        self.h = lib().vdo_mg_create(kind, m, n, cx, cy, x_sym)
        lib().vdo_mg_set(self.h, nu1, nu2, ncs)

    def close(self):
        if self.h:
            lib().vdo_mg_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def levels(self):
        return lib().vdo_mg_levels(self.h)

    def apply(self, z):
        z = np.ascontiguousarray(z, np.float64).ravel()
        out = np.empty_like(z)
        lib().vdo_mg_apply(self.h, ptr(z), ptr(out))
        return out

    def resnorm(self, b, z):
        b = np.ascontiguousarray(b, np.float64).ravel()
        z = np.ascontiguousarray(z, np.float64).ravel()
        return lib().vdo_mg_resnorm(self.h, ptr(b), ptr(z))

    def solve(self, b, z0=None, tol=1e-10, max_cycles=200):
        b = np.ascontiguousarray(b, np.float64).ravel()
        z = np.zeros_like(b) if z0 is None else np.array(z0, np.float64).ravel()
        c = lib().vdo_mg_solve(self.h, ptr(b), ptr(z), tol, max_cycles)
        return z, c

    def vcycles(self, b, z0, k):
        b = np.ascontiguousarray(b, np.float64).ravel()
        z = np.array(z0, np.float64).ravel()
        lib().vdo_mg_vcycles(self.h, ptr(b), ptr(z), k)
        return z
